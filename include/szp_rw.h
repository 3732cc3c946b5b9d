/*
 * szp_rw.h -- the raw-file helpers of the reference's I/O module that its
 * example drivers use, provided by libszp_amd.so so that code written
 * against SZp's API links unchanged.  Each declaration names the reference
 * interface it replaces (SZp/include/szp_rw.h, SZp/src/szp_rw.cc).
 *
 * Files are raw arrays with no header: a stream file holds the bytes of
 * szp_compress(), a data file the values.  Values are read in the byte order
 * szp_dataEndianType names and converted to the host's (szp_sysEndianType),
 * as the reference does.  status receives SZ_SCES (0) or SZ_FERR (-2).
 */
#ifndef SZP_AMD_SZP_RW_H
#define SZP_AMD_SZP_RW_H

#include <stddef.h>

/* Status codes of SZp/include/szp_defines.h:59-61. */
#define SZ_SCES 0
#define SZ_NSCS -1
#define SZ_FERR -2

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces szp_checkFileSize, SZp/include/szp_rw.h:37. */
size_t szp_checkFileSize(char *srcFilePath, int *status);

/* Replaces szp_readByteData, SZp/include/szp_rw.h:39 (SZp/src/szp_rw.cc:135-161):
 * the whole file, malloc'd (caller free()s it). */
unsigned char *szp_readByteData(char *srcFilePath, size_t *byteLength, int *status);

/* Replace szp_readFloatData / szp_readDoubleData, SZp/include/szp_rw.h:48,40
 * (SZp/src/szp_rw.cc:163-198,424-459): *nbEle = file bytes / value size. */
float *szp_readFloatData(char *srcFilePath, size_t *nbEle, int *status);
double *szp_readDoubleData(char *srcFilePath, size_t *nbEle, int *status);

/* Replaces szp_writeByteData, SZp/include/szp_rw.h:61 (SZp/src/szp_rw.cc:781-795). */
void szp_writeByteData(unsigned char *bytes, size_t byteLength, char *tgtFilePath, int *status);

/* Replace szp_writeFloatData_inBytes / szp_writeDoubleData_inBytes,
 * SZp/include/szp_rw.h:65-66 (SZp/src/szp_rw.cc:865-906): the values'
 * host-order bytes. */
void szp_writeFloatData_inBytes(float *data, size_t nbEle, char *tgtFilePath, int *status);
void szp_writeDoubleData_inBytes(double *data, size_t nbEle, char *tgtFilePath, int *status);

#ifdef __cplusplus
}
#endif

#endif /* SZP_AMD_SZP_RW_H */
