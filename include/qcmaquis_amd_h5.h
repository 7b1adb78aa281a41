/* Result and checkpoint files of the DMRG driver (HDF5), host only: libqcmaquis_amd_h5.so.
 *
 * The reference writes its results through ALPS archives (applications/dmrg/dmrg_sim.hpp:88-100:
 * /spectrum/iteration/<sweep>/results/<name>/mean/value, /spectrum/results/Energy/mean/value;
 * utils/results_collector.h:56-63 appends to "mean/value") and the checkpoint status to
 * <chkp>/props.h5:/status (sim/sim.hpp:223-235).  These calls read and write such datasets:
 * doubles (F64LE), unsigned 64-bit (U64LE), 32-bit integers (I32LE) and variable-length ASCII
 * strings, scalar (rank 0) or one-dimensional (rank 1), the types the reference's files hold. */
#ifndef QCMAQUIS_AMD_H5_H
#define QCMAQUIS_AMD_H5_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MQ_H5_OK           0
#define MQ_H5_ERR_INVALID -1  /* bad argument, or the output buffer is too small */
#define MQ_H5_ERR_IO      -2  /* the file cannot be opened / created, or an HDF5 call failed */
#define MQ_H5_ERR_MISSING -3  /* no such dataset or group */
#define MQ_H5_ERR_TYPE    -4  /* the dataset has another type */

#define MQ_H5_FLOAT   1
#define MQ_H5_INTEGER 2
#define MQ_H5_STRING  3

/* write (create the file and groups as needed; replace an existing dataset) */
int mq_h5_write_f64(const char* file, const char* path, const double* v, uint64_t n, int rank);
int mq_h5_write_u64(const char* file, const char* path, const uint64_t* v, uint64_t n, int rank);
int mq_h5_write_i32(const char* file, const char* path, const int32_t* v, uint64_t n, int rank);
int mq_h5_write_str(const char* file, const char* path, const char* s);
/* 1 if the dataset or group exists, 0 if not */
int mq_h5_exists(const char* file, const char* path);
/* type class (MQ_H5_FLOAT / INTEGER / STRING), rank and element count of a dataset */
int mq_h5_info(const char* file, const char* path, int* type_class, int* rank, uint64_t* n);
/* read a float or integer dataset as doubles, an integer dataset as int64, a string dataset */
int mq_h5_read_f64(const char* file, const char* path, double* out, uint64_t cap);
int mq_h5_read_i64(const char* file, const char* path, int64_t* out, uint64_t cap);
int mq_h5_read_str(const char* file, const char* path, char* out, uint64_t cap);
/* member names of a group, separated by '\n' */
int mq_h5_list(const char* file, const char* group, char* out, uint64_t cap);

#ifdef __cplusplus
}
#endif
#endif
