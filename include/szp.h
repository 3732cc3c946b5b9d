/*
 * szp.h -- drop-in C ABI for SZp's compress/decompress (fastMode 1 and 2,
 * float32 and float64), backed by MI355X HIP kernels (libszp_amd.so).
 *
 * Each declaration keeps the reference's name, signature, ownership and
 * error convention; the comment names the reference interface it replaces
 * (paths relative to the SZp repository root).  fastMode SZP_NONRANDOMACCESS
 * and SZP_RANDOMACCESS with dataType SZ_FLOAT or SZ_DOUBLE are served; other
 * modes return NULL (see szp_amd_last_error()).
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
 * blocks, not from byteLength, and every fastMode other than 1 decodes as
 * random access.  Returns a malloc'd float (SZ_FLOAT) or double (SZ_DOUBLE)
 * array (caller free()s it) or NULL. */
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

/* Replaces szp_float_single_thread_arg, SZp/include/szp_float.h:36-37
 * (SZp/src/szp_float.cc:482-605): the fastMode-1 stream of a team of one
 * (T = 1, whatever szp_amd_default_threads() says).  `output` must hold
 * szp_amd_compress_bound(nbEle, blockSize, 1) bytes. */
void szp_float_single_thread_arg(unsigned char *output, float *oriData, size_t *outSize,
                                 float absErrBound, size_t nbEle, int blockSize);

/* Replaces szp_float_single_thread_arg_record, SZp/include/szp_float.h:39-40
 * (SZp/src/szp_float.cc:607-750): as szp_float_single_thread_arg, returning
 * the memory-access count the reference's instrumented loop tallies for the
 * same stream (0 on failure). */
size_t szp_float_single_thread_arg_record(unsigned char *output, float *oriData, size_t *outSize,
                                          float absErrBound, size_t nbEle, int blockSize);

/* Replaces szp_float_decompress_single_thread_arg, SZp/include/szpd_float.h:25
 * (SZp/src/szpd_float.cc:130-214): decodes a T = 1 stream; cmpBytes points
 * just after the 4-byte eb. */
void szp_float_decompress_single_thread_arg(float *newData, size_t nbEle, float absErrBound,
                                            int blockSize, unsigned char *cmpBytes);

/* Replaces szp_float_decompress_single_thread_arg_record, SZp/include/szpd_float.h:27
 * (SZp/src/szpd_float.cc:216-317): as szp_float_decompress_single_thread_arg,
 * returning the reference's memory-access count (0 on failure). */
size_t szp_float_decompress_single_thread_arg_record(float *newData, size_t nbEle,
                                                     float absErrBound, int blockSize,
                                                     unsigned char *cmpBytes);

/* Replaces szp_float_openmp_threadblock_randomaccess, SZp/include/szp_float.h:42-44
 * (SZp/src/szp_float.cc:752-912): fastMode 2, each chunk cut into units of
 * blockSize values, a unit = the raw quantized first value and one block of
 * the others. */
unsigned char *szp_float_openmp_threadblock_randomaccess(float *oriData, size_t *outSize,
                                                         float absErrBound, size_t nbEle,
                                                         int blockSize);

/* Replaces szp_float_openmp_threadblock_randomaccess_arg, SZp/include/szp_float.h:46-48.
 * `output` must hold szp_amd_compress_bound_typed(SZP_RANDOMACCESS, SZ_FLOAT,
 * nbEle, blockSize, T) bytes. */
void szp_float_openmp_threadblock_randomaccess_arg(unsigned char *output, float *oriData,
                                                   size_t *outSize, float absErrBound,
                                                   size_t nbEle, int blockSize);

/* Replaces szp_float_decompress_openmp_threadblock_randomaccess,
 * SZp/include/szpd_float.h:29 (SZp/src/szpd_float.cc:425-530). */
float *szp_float_decompress_openmp_threadblock_randomaccess(size_t nbEle, float absErrBound,
                                                            int blockSize, unsigned char *cmpBytes);

/* Replaces szp_float_decompress_openmp_threadblock_randomaccess_arg,
 * SZp/include/szpd_float.h:31. */
void szp_float_decompress_openmp_threadblock_randomaccess_arg(float *newData, size_t nbEle,
                                                              float absErrBound, int blockSize,
                                                              unsigned char *cmpBytes);

/* Replaces szp_double_openmp_threadblock, SZp/include/szp_double.h:30-31
 * (SZp/src/szp_double.cc:148-318): the float format with an 8-byte
 * big-endian eb and the quantizer (int)(x * (1 / eb)) in double. */
unsigned char *szp_double_openmp_threadblock(double *oriData, size_t *outSize, double absErrBound,
                                             size_t nbEle, int blockSize);

/* Replaces szp_double_openmp_threadblock_arg, SZp/include/szp_double.h:33-34.
 * `output` must hold szp_amd_compress_bound_typed(SZ_DOUBLE, nbEle,
 * blockSize, T) bytes. */
void szp_double_openmp_threadblock_arg(unsigned char *output, double *oriData, size_t *outSize,
                                       double absErrBound, size_t nbEle, int blockSize);

/* Replaces szp_double_decompress_openmp_threadblock, SZp/include/szpd_double.h:21
 * (SZp/src/szpd_double.cc:23-106).  cmpBytes points just after the 8-byte eb. */
double *szp_double_decompress_openmp_threadblock(size_t nbEle, double absErrBound, int blockSize,
                                                 unsigned char *cmpBytes);

/* Replaces szp_double_decompress_openmp_threadblock_arg, SZp/include/szpd_double.h:23. */
void szp_double_decompress_openmp_threadblock_arg(double *newData, size_t nbEle,
                                                  double absErrBound, int blockSize,
                                                  unsigned char *cmpBytes);

/* Replaces szp_double_openmp_threadblock_randomaccess, SZp/include/szp_double.h:42-44
 * (SZp/src/szp_double.cc:759-920). */
unsigned char *szp_double_openmp_threadblock_randomaccess(double *oriData, size_t *outSize,
                                                          double absErrBound, size_t nbEle,
                                                          int blockSize);

/* Replaces szp_double_openmp_threadblock_randomaccess_arg, SZp/include/szp_double.h:46-48. */
void szp_double_openmp_threadblock_randomaccess_arg(unsigned char *output, double *oriData,
                                                    size_t *outSize, double absErrBound,
                                                    size_t nbEle, int blockSize);

/* Replaces szp_double_decompress_openmp_threadblock_randomaccess,
 * SZp/include/szpd_double.h:29 (SZp/src/szpd_double.cc:431-534). */
double *szp_double_decompress_openmp_threadblock_randomaccess(size_t nbEle, double absErrBound,
                                                              int blockSize,
                                                              unsigned char *cmpBytes);

/* Replaces szp_double_decompress_openmp_threadblock_randomaccess_arg,
 * SZp/include/szpd_double.h:31. */
void szp_double_decompress_openmp_threadblock_randomaccess_arg(double *newData, size_t nbEle,
                                                               double absErrBound, int blockSize,
                                                               unsigned char *cmpBytes);

#ifdef __cplusplus
}
#endif

#endif /* SZP_AMD_SZP_H */
