/*
 * szp.h -- drop-in C ABI for SZp's fastmode1 float32 path, backed by MI355X
 * HIP kernels (libszp_amd.so).
 *
 * Each declaration keeps the reference's name, signature, ownership and
 * error convention; the comment names the reference interface it replaces
 * (paths relative to the SZp repository root).  Only fastMode ==
 * SZP_NONRANDOMACCESS with dataType == SZ_FLOAT is served: other modes return
 * NULL (see szp_amd_last_error()).
 *
 * T, the reference's OpenMP team size, is part of the stream format
 * (SZp/src/szp_float.cc:179-181).  These entry points take it from
 * szp_amd_default_threads(): szp_amd_set_threads() if called, else
 * SZP_AMD_NUM_THREADS, else omp_get_max_threads() of an OpenMP runtime already
 * loaded in the process, else OMP_NUM_THREADS, else the CPUs this process may
 * run on -- the value the reference's own parallel region would use.
 */
#ifndef SZP_AMD_SZP_H
#define SZP_AMD_SZP_H

#include <stddef.h>

/* Constants of SZp/include/szp_defines.h:19-20,33-37. */
#define ABS 0
#define REL 1
#define SZP_NONRANDOMACCESS 1
#define SZP_RANDOMACCESS 2
#define SZ_FLOAT 0
#define SZ_DOUBLE 1
#define LITTLE_ENDIAN_DATA 0
#define BIG_ENDIAN_DATA 1
#define LITTLE_ENDIAN_SYSTEM 0
#define BIG_ENDIAN_SYSTEM 1

#ifdef __cplusplus
extern "C" {
#endif

/* Exported globals of SZp/src/szp.cc:17-20 (declared SZp/include/szp_defines.h:103-107). */
extern int szp_versionNumber[4];
extern int szp_dataEndianType;
extern int szp_sysEndianType;

/* Replaces szp_compress, SZp/include/szp.h:33-34 (SZp/src/szp.cc:23-115).
 * Returns a malloc'd stream the caller releases with free(); *outSize gets
 * its length.  NULL on an unsupported mode or a device error. */
unsigned char *szp_compress(int fastMode, int dataType, void *data, size_t *outSize,
                            int errBoundMode, double absErrBound, double relBoundRatio,
                            size_t nbEle, int blockSize);

/* Replaces szp_decompress, SZp/include/szp.h:36 (SZp/src/szp.cc:118-179).
 * Like the reference, the stream's extent comes from its own header and
 * blocks, not from byteLength.  Returns a malloc'd float array (caller
 * free()s it) or NULL. */
void *szp_decompress(int fastMode, int dataType, unsigned char *bytes, size_t byteLength,
                     size_t nbEle, int blockSize);

/* Replaces szp_float_openmp_threadblock, SZp/include/szp_float.h:29-31
 * (SZp/src/szp_float.cc:148-309). */
unsigned char *szp_float_openmp_threadblock(float *oriData, size_t *outSize, float absErrBound,
                                            size_t nbEle, int blockSize);

/* Replaces szp_float_openmp_threadblock_arg, SZp/include/szp_float.h:33-34
 * (SZp/src/szp_float.cc:314-480).  `output` must hold
 * szp_amd_compress_bound(nbEle, blockSize, T) bytes. */
void szp_float_openmp_threadblock_arg(unsigned char *output, float *oriData, size_t *outSize,
                                      float absErrBound, size_t nbEle, int blockSize);

/* Replaces szp_float_decompress_openmp_threadblock, SZp/include/szpd_float.h:21
 * (SZp/src/szpd_float.cc:23-128).  cmpBytes points just after the 4-byte eb. */
float *szp_float_decompress_openmp_threadblock(size_t nbEle, float absErrBound, int blockSize,
                                               unsigned char *cmpBytes);

/* Replaces szp_float_decompress_openmp_threadblock_arg, SZp/include/szpd_float.h:23
 * (SZp/src/szpd_float.cc:320-423). */
void szp_float_decompress_openmp_threadblock_arg(float *newData, size_t nbEle, float absErrBound,
                                                 int blockSize, unsigned char *cmpBytes);

#ifdef __cplusplus
}
#endif

#endif /* SZP_AMD_SZP_H */
