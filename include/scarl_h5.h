/* scarl_h5.h -- C ABI of libscarl_h5.so: the HDF5 and XDMF output of a ScaRL sample.
 *
 * Replaces the writers of the reference's `write_output` and `hdf5_ScaRL` modules
 * (SRC_F/write_output.f90, SRC_F/hdf5_ScaRL.f90), which go through the HDF5 Fortran API. This
 * library uses the HDF5 C API, so the files have the same on-disk layout:
 *
 *   - A field is a rank-3 dataset whose Fortran extents are (n1, n2, n3), x fastest. HDF5
 *     records such a dataset with the extents reversed, {n3, n2, n1}. That is the layout
 *     h5screate_simple_f gives a Fortran array.
 *   - Offsets and counts in this header are Fortran-ordered (x first) and 0-based.
 *   - Vector attributes are 1D; scalar attributes use a scalar dataspace.
 *   - Label strings are Fortran strings: space padded, SCARL_H5_LABEL_LEN characters each.
 *
 * Precision codes: 64 = H5T_NATIVE_DOUBLE (the reference's type), 32 = H5T_NATIVE_FLOAT.
 * Every entry point returns 0 on success or a SCARL_H5_E_* code. The message of the last
 * failure on the calling thread is available from scarl_h5_last_error.
 */
#ifndef SCARL_H5_H
#define SCARL_H5_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { SCARL_H5_OK = 0, SCARL_H5_E_INVALID = 1, SCARL_H5_E_IO = 2 };
#define SCARL_H5_LABEL_LEN 20 /* character(len=20) time labels, SRC_F/main_ScaRL.f90:40 */

typedef struct scarl_h5 scarl_h5; /* one open HDF5 file */

/* Message of the last failure on this thread (NUL terminated, truncated to len). */
int scarl_h5_last_error(char* buf, int32_t len);

/* h5fcreate_f(path, H5F_ACC_TRUNC_F): write_output.f90:379 (also :80, :226). */
int scarl_h5_create(const char* path, scarl_h5** out);
/* h5fopen_f(path, H5F_ACC_RDWR_F): write_output.f90:747, :882; writable = 0 opens read-only
 * (the Stat program's readers). */
int scarl_h5_open(const char* path, int32_t writable, scarl_h5** out);
int scarl_h5_close(scarl_h5* f);

/* h5screate_simple_f + h5dcreate_f of a rank-3 dataset of Fortran extents dims
 * (write_output.f90:380-382). file_prec: 64 (H5T_NATIVE_DOUBLE) or 32. */
int scarl_h5_create_dataset(scarl_h5* f, const char* name, const int32_t dims[3], int32_t file_prec);

/* h5sselect_hyperslab_f + h5dwrite_f (write_output.f90:383-391, :414-423): writes a block of
 * Fortran extents count at 0-based offset. Memory is x fastest with element strides ld1
 * (between y rows) and ld2 (between z planes); ld1 >= count[0], ld2 >= ld1 * count[1], and
 * ld2 is a multiple of ld1. mem_prec: 64 or 32. HDF5 converts it to the dataset type. */
int scarl_h5_write_block(scarl_h5* f, const char* ds_name, const void* data, int32_t mem_prec, int64_t ld1,
                         int64_t ld2, const int32_t offset[3], const int32_t count[3]);

/* Attribute helpers of hdf5_ScaRL.f90. object = "" attaches to the file root
 * (write_HDF5_attributes, write_output.f90:722-813), otherwise to the named dataset
 * (write_output.f90:393, :426, :700-716). n = 0 writes a scalar (write_h5attr_int / _real,
 * hdf5_ScaRL.f90:31-72); n >= 1 writes a vector (write_h5attr_int_vec / _real_vec, :105-172).
 * Integers are stored as H5T_STD_I32LE and doubles as H5T_NATIVE_DOUBLE. */
int scarl_h5_attr_i32(scarl_h5* f, const char* object, const char* name, const int32_t* v, int32_t n);
int scarl_h5_attr_f64(scarl_h5* f, const char* object, const char* name, const double* v, int32_t n);

/* write_time_on_hdf5 (write_output.f90:870-902): dataset "times" of Fortran extents
 * (ntimes, nprocs) (H5T_NATIVE_DOUBLE, column-major times), and dataset "time_labels" of
 * ntimes Fortran strings (H5T_FORTRAN_S1 of size SCARL_H5_LABEL_LEN). labels holds
 * ntimes * SCARL_H5_LABEL_LEN characters. */
int scarl_h5_write_times(scarl_h5* f, const char* labels, const double* times, int32_t ntimes, int32_t nprocs);

/* Readers: the Stat program re-reads these files (SRC_F/Stat_ScaRL/hdf5_RF.f90).
 * scarl_h5_dataset_info returns the Fortran extents (x first; trailing extents 1 for a dataset
 * of rank < 3) and the stored precision (64 or 32). */
int scarl_h5_dataset_info(scarl_h5* f, const char* name, int32_t dims[3], int32_t* file_prec);
int scarl_h5_read_block(scarl_h5* f, const char* ds_name, void* data, int32_t mem_prec, int64_t ld1, int64_t ld2,
                        const int32_t offset[3], const int32_t count[3]);
/* Reads up to cap values of a scalar or 1D attribute; *n receives its length (1 for a scalar). */
int scarl_h5_read_attr_i32(scarl_h5* f, const char* object, const char* name, int32_t* v, int32_t cap, int32_t* n);
int scarl_h5_read_attr_f64(scarl_h5* f, const char* object, const char* name, double* v, int32_t cap, int32_t* n);
/* Number of attributes of the object and the name of the idx-th (creation order not tracked:
 * HDF5's name order). */
int scarl_h5_attr_count(scarl_h5* f, const char* object, int32_t* n);
int scarl_h5_attr_name(scarl_h5* f, const char* object, int32_t idx, char* buf, int32_t len);
/* Reads "times" and "time_labels". Up to cap_times values and cap_labels labels are copied;
 * *ntimes and *nprocs receive the extents. */
int scarl_h5_read_times(scarl_h5* f, double* times, int32_t cap_times, char* labels, int32_t cap_labels,
                        int32_t* ntimes, int32_t* nprocs);

/* write_XMF_elements (write_output.f90:455-556): the .xmf of one dataset on a 3DCoRectMesh.
 * Extents are written in reverse order. The origin and step are xMin and
 * (xMax - xMin) / (L - 1), also reversed, in Fortran F25.10 records. The file is
 * xmf_folder/xmf_name, and the data item points at h5_to_xmf_path/h5_name:/ds_name.
 * file_prec 32 writes Precision="4" for an FP32 dataset; the reference always writes "8". */
int scarl_xmf_write_elements(const char* xmf_folder, const char* xmf_name, const char* h5_name,
                             const char* h5_to_xmf_path, const char* ds_name, const double xMin[3],
                             const double xMax[3], const int32_t L[3], int32_t file_prec);

/* write_XMF_global (write_output.f90:558-656): one grid per subdomain i < nb_procs. The grid's
 * data item is h5_list[i]:/sample_1_p<i>, its origin is coord_0_list[3 i .. 3 i + 2], and the
 * spacing xStep and extents xNStep are shared. The file is path (the reference opens
 * "./" // globXMF_name, globXMF_name = res_folder/output_name.xmf, main_ScaRL.f90:664). */
int scarl_xmf_write_global(const char* path, const char* const* h5_list, const double* coord_0_list,
                           const double xStep[3], const int32_t xNStep[3], int32_t nb_procs);

/* Fortran Fw.d edit of one value (gfortran conventions: leading zero where it fits, "Infinity",
 * "NaN", w asterisks when the value does not fit). Writes w characters and a NUL into buf
 * (len > w). Exposed for the tests of the .xmf records. */
int scarl_fortran_fixed(double v, int32_t w, int32_t d, char* buf, int32_t len);

#ifdef __cplusplus
}
#endif

#endif /* SCARL_H5_H */
