// scarl_h5.cpp -- HDF5 and XDMF output of a ScaRL sample through the HDF5 C API (libscarl_h5.so).
//
// The on-disk layout is that of the reference's Fortran writers (SRC_F/write_output.f90,
// SRC_F/hdf5_ScaRL.f90). The HDF5 Fortran API stores an array of Fortran extents (n1, n2, n3)
// as a dataspace {n3, n2, n1}, with offsets reversed the same way, so every rank-3 extent and
// offset here is reversed at the boundary. Integer attributes are H5T_STD_I32LE and real ones
// H5T_NATIVE_DOUBLE, as in hdf5_ScaRL.f90. The .xmf files are written record by record, as the
// Fortran formatted WRITE statements of write_output.f90 produce them.
#include "../../include/scarl_h5.h"

#include <hdf5.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

struct scarl_h5 {
    hid_t file;
};

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

// HDF5's own error stack printing is switched off once; failures come back as codes.
void quiet() {
    static std::once_flag once;
    std::call_once(once, [] {
        H5open();
        H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    });
}

// Closes an identifier at scope exit.
struct Hid {
    hid_t id = -1;
    herr_t (*close)(hid_t) = nullptr;
    Hid(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
    ~Hid() {
        if (id >= 0 && close) close(id);
    }
    Hid(const Hid&) = delete;
    Hid& operator=(const Hid&) = delete;
    bool ok() const { return id >= 0; }
};

hid_t mem_type(int32_t prec) {
    return prec == 64 ? H5T_NATIVE_DOUBLE : prec == 32 ? H5T_NATIVE_FLOAT : -1;
}

// The object an attribute goes on: the file root ("" or "/") or a dataset.
hid_t open_object(scarl_h5* f, const char* object) {
    if (!object || !*object || !std::strcmp(object, "/")) return H5Oopen(f->file, "/", H5P_DEFAULT);
    return H5Oopen(f->file, object, H5P_DEFAULT);
}

bool valid_block(const int32_t offset[3], const int32_t count[3], int64_t ld1, int64_t ld2) {
    for (int d = 0; d < 3; ++d)
        if (offset[d] < 0 || count[d] < 1) return false;
    return ld1 >= count[0] && ld2 >= ld1 * count[1] && ld2 % ld1 == 0;
}

// Memory dataspace {count3, ld2 / ld1, ld1} with the block {count3, count2, count1} selected.
hid_t block_memspace(const int32_t count[3], int64_t ld1, int64_t ld2) {
    const hsize_t dims[3] = {(hsize_t)count[2], (hsize_t)(ld2 / ld1), (hsize_t)ld1};
    hid_t s = H5Screate_simple(3, dims, nullptr);
    if (s < 0) return s;
    const hsize_t start[3] = {0, 0, 0}, cnt[3] = {(hsize_t)count[2], (hsize_t)count[1], (hsize_t)count[0]};
    if (H5Sselect_hyperslab(s, H5S_SELECT_SET, start, nullptr, cnt, nullptr) < 0) {
        H5Sclose(s);
        return -1;
    }
    return s;
}

// A rank-3 dataset's file space with the block selected. The dataset extents come back
// Fortran-ordered in fdims.
hid_t block_filespace(hid_t dset, const int32_t offset[3], const int32_t count[3], hsize_t fdims[3], std::string& why) {
    hid_t s = H5Dget_space(dset);
    if (s < 0) {
        why = "no dataspace";
        return -1;
    }
    if (H5Sget_simple_extent_ndims(s) != 3) {
        H5Sclose(s);
        why = "dataset is not rank 3";
        return -1;
    }
    hsize_t d[3];
    H5Sget_simple_extent_dims(s, d, nullptr);
    fdims[0] = d[2];
    fdims[1] = d[1];
    fdims[2] = d[0];
    for (int a = 0; a < 3; ++a)
        if ((hsize_t)offset[a] + (hsize_t)count[a] > fdims[a]) {
            H5Sclose(s);
            why = "block exceeds the dataset extents";
            return -1;
        }
    const hsize_t start[3] = {(hsize_t)offset[2], (hsize_t)offset[1], (hsize_t)offset[0]};
    const hsize_t cnt[3] = {(hsize_t)count[2], (hsize_t)count[1], (hsize_t)count[0]};
    // h5sselect_none_f then H5S_SELECT_SET, as write_output.f90:385-386
    H5Sselect_none(s);
    if (H5Sselect_hyperslab(s, H5S_SELECT_SET, start, nullptr, cnt, nullptr) < 0) {
        H5Sclose(s);
        why = "hyperslab selection failed";
        return -1;
    }
    return s;
}

int write_attr(scarl_h5* f, const char* object, const char* name, hid_t ftype, hid_t mtype, const void* v,
               int32_t n) {
    quiet();
    if (!f || !name || !*name || !v || n < 0) return fail(SCARL_H5_E_INVALID, "attribute: invalid argument");
    Hid obj(open_object(f, object), H5Oclose);
    if (!obj.ok()) return fail(SCARL_H5_E_IO, "attribute %s: cannot open object '%s'", name, object ? object : "");
    hsize_t dims[1] = {(hsize_t)n};
    Hid space(n == 0 ? H5Screate(H5S_SCALAR) : H5Screate_simple(1, dims, dims), H5Sclose);
    if (!space.ok()) return fail(SCARL_H5_E_IO, "attribute %s: dataspace", name);
    if (H5Aexists(obj.id, name) > 0) H5Adelete(obj.id, name);
    Hid attr(H5Acreate2(obj.id, name, ftype, space.id, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
    if (!attr.ok()) return fail(SCARL_H5_E_IO, "attribute %s: create failed", name);
    if (H5Awrite(attr.id, mtype, v) < 0) return fail(SCARL_H5_E_IO, "attribute %s: write failed", name);
    return SCARL_H5_OK;
}

int read_attr(scarl_h5* f, const char* object, const char* name, hid_t mtype, void* v, size_t esize, int32_t cap,
              int32_t* n) {
    quiet();
    if (!f || !name || !n || cap < 0 || (cap > 0 && !v)) return fail(SCARL_H5_E_INVALID, "attribute read: invalid argument");
    Hid obj(open_object(f, object), H5Oclose);
    if (!obj.ok()) return fail(SCARL_H5_E_IO, "attribute %s: cannot open object '%s'", name, object ? object : "");
    if (H5Aexists(obj.id, name) <= 0) return fail(SCARL_H5_E_IO, "attribute %s: not found", name);
    Hid attr(H5Aopen(obj.id, name, H5P_DEFAULT), H5Aclose);
    Hid space(H5Aget_space(attr.id), H5Sclose);
    if (!attr.ok() || !space.ok()) return fail(SCARL_H5_E_IO, "attribute %s: open failed", name);
    const hssize_t np = H5Sget_simple_extent_npoints(space.id);
    if (np < 1) return fail(SCARL_H5_E_IO, "attribute %s: empty", name);
    *n = (int32_t)np;
    std::vector<unsigned char> tmp((size_t)np * esize);
    if (H5Aread(attr.id, mtype, tmp.data()) < 0) return fail(SCARL_H5_E_IO, "attribute %s: read failed", name);
    std::memcpy(v, tmp.data(), (size_t)std::min<hssize_t>(np, cap) * esize);
    return SCARL_H5_OK;
}

// ---- Fortran formatted output ---------------------------------------------------------------

// Fw.d of one value (gfortran): w characters, right justified.
std::string fixed(double v, int w, int d) {
    std::string s;
    if (std::isnan(v)) {
        s = "NaN";
    } else if (std::isinf(v)) {
        s = w >= 9 || (w == 8 && v > 0) ? (v < 0 ? "-Infinity" : "Infinity") : (v < 0 ? "-Inf" : "Inf");
    } else {
        char buf[512];
        std::snprintf(buf, sizeof buf, "%.*f", d, v);
        s = buf;
    }
    if ((int)s.size() > w) return std::string((size_t)w, '*');
    return std::string((size_t)w - s.size(), ' ') + s;
}

// num2str (SRC_F/str_functions.f90:152-194) without nCharacters: the I30 text, left adjusted,
// in a character(len=30) result. trim() of it is the plain decimal.
std::string num2str30(long v) {
    std::string s = std::to_string(v);
    return s + std::string(30 - s.size(), ' ');
}

// dimText of write_XMF_elements / write_XMF_global: the extents in reverse order, blank separated.
std::string dim_text(const int32_t n[3]) {
    return std::to_string(n[2]) + " " + std::to_string(n[1]) + " " + std::to_string(n[0]);
}

// The records of WRITE(file,'(A,F25.10)') ' ', a, ' ', b, ' ', c: the format is used up after
// two items and reverts, so each value is its own record " " // F25.10.
void f25_records(std::string& out, double a, double b, double c) {
    for (double v : {a, b, c}) out += " " + fixed(v, 25, 10) + "\n";
}

int write_text(const std::string& path, const std::string& text) {
    FILE* fp = std::fopen(path.c_str(), "w");
    if (!fp) return fail(SCARL_H5_E_IO, "cannot open %s for writing", path.c_str());
    const size_t w = std::fwrite(text.data(), 1, text.size(), fp);
    const int rc = std::fclose(fp);
    if (w != text.size() || rc != 0) return fail(SCARL_H5_E_IO, "short write to %s", path.c_str());
    return SCARL_H5_OK;
}

std::string join_path(const char* folder, const char* name) {
    return std::string(folder && *folder ? folder : ".") + "/" + name;
}

}  // namespace

extern "C" {

int scarl_h5_last_error(char* buf, int32_t len) {
    if (!buf || len <= 0) return SCARL_H5_E_INVALID;
    std::snprintf(buf, (size_t)len, "%s", g_err.c_str());
    return SCARL_H5_OK;
}

int scarl_h5_create(const char* path, scarl_h5** out) {
    quiet();
    if (!path || !*path || !out) return fail(SCARL_H5_E_INVALID, "create: invalid argument");
    const hid_t fid = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (fid < 0) return fail(SCARL_H5_E_IO, "cannot create %s", path);
    *out = new scarl_h5{fid};
    return SCARL_H5_OK;
}

int scarl_h5_open(const char* path, int32_t writable, scarl_h5** out) {
    quiet();
    if (!path || !*path || !out) return fail(SCARL_H5_E_INVALID, "open: invalid argument");
    const hid_t fid = H5Fopen(path, writable ? H5F_ACC_RDWR : H5F_ACC_RDONLY, H5P_DEFAULT);
    if (fid < 0) return fail(SCARL_H5_E_IO, "cannot open %s", path);
    *out = new scarl_h5{fid};
    return SCARL_H5_OK;
}

int scarl_h5_close(scarl_h5* f) {
    if (!f) return fail(SCARL_H5_E_INVALID, "close: null file");
    const herr_t rc = H5Fclose(f->file);
    delete f;
    return rc < 0 ? fail(SCARL_H5_E_IO, "close failed") : SCARL_H5_OK;
}

int scarl_h5_create_dataset(scarl_h5* f, const char* name, const int32_t dims[3], int32_t file_prec) {
    quiet();
    if (!f || !name || !*name || !dims) return fail(SCARL_H5_E_INVALID, "create_dataset: invalid argument");
    const hid_t ftype = mem_type(file_prec);
    if (ftype < 0) return fail(SCARL_H5_E_INVALID, "create_dataset: precision %d (64 or 32)", file_prec);
    for (int d = 0; d < 3; ++d)
        if (dims[d] < 1) return fail(SCARL_H5_E_INVALID, "create_dataset: extent %d < 1", dims[d]);
    const hsize_t ds[3] = {(hsize_t)dims[2], (hsize_t)dims[1], (hsize_t)dims[0]};
    Hid space(H5Screate_simple(3, ds, nullptr), H5Sclose);
    if (!space.ok()) return fail(SCARL_H5_E_IO, "create_dataset %s: dataspace", name);
    Hid dset(H5Dcreate2(f->file, name, ftype, space.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
    if (!dset.ok()) return fail(SCARL_H5_E_IO, "create_dataset %s failed (exists?)", name);
    return SCARL_H5_OK;
}

int scarl_h5_write_block(scarl_h5* f, const char* ds_name, const void* data, int32_t mem_prec, int64_t ld1,
                         int64_t ld2, const int32_t offset[3], const int32_t count[3]) {
    quiet();
    if (!f || !ds_name || !data || !offset || !count) return fail(SCARL_H5_E_INVALID, "write_block: invalid argument");
    const hid_t mtype = mem_type(mem_prec);
    if (mtype < 0) return fail(SCARL_H5_E_INVALID, "write_block: precision %d (64 or 32)", mem_prec);
    if (!valid_block(offset, count, ld1, ld2))
        return fail(SCARL_H5_E_INVALID, "write_block: block or strides invalid (ld1 %lld ld2 %lld)", (long long)ld1,
                    (long long)ld2);
    Hid dset(H5Dopen2(f->file, ds_name, H5P_DEFAULT), H5Dclose);
    if (!dset.ok()) return fail(SCARL_H5_E_IO, "write_block: no dataset %s", ds_name);
    hsize_t fd[3];
    std::string why;
    Hid fs(block_filespace(dset.id, offset, count, fd, why), H5Sclose);
    if (!fs.ok()) return fail(SCARL_H5_E_INVALID, "write_block %s: %s", ds_name, why.c_str());
    Hid ms(block_memspace(count, ld1, ld2), H5Sclose);
    if (!ms.ok()) return fail(SCARL_H5_E_IO, "write_block %s: memory dataspace", ds_name);
    if (H5Dwrite(dset.id, mtype, ms.id, fs.id, H5P_DEFAULT, data) < 0)
        return fail(SCARL_H5_E_IO, "write_block %s: H5Dwrite failed", ds_name);
    return SCARL_H5_OK;
}

int scarl_h5_attr_i32(scarl_h5* f, const char* object, const char* name, const int32_t* v, int32_t n) {
    return write_attr(f, object, name, H5T_STD_I32LE, H5T_NATIVE_INT32, v, n);
}

int scarl_h5_attr_f64(scarl_h5* f, const char* object, const char* name, const double* v, int32_t n) {
    return write_attr(f, object, name, H5T_NATIVE_DOUBLE, H5T_NATIVE_DOUBLE, v, n);
}

int scarl_h5_write_times(scarl_h5* f, const char* labels, const double* times, int32_t ntimes, int32_t nprocs) {
    quiet();
    if (!f || !labels || !times || ntimes < 1 || nprocs < 1) return fail(SCARL_H5_E_INVALID, "write_times: invalid argument");
    {
        const hsize_t ds[2] = {(hsize_t)nprocs, (hsize_t)ntimes};
        Hid space(H5Screate_simple(2, ds, nullptr), H5Sclose);
        Hid dset(H5Dcreate2(f->file, "times", H5T_NATIVE_DOUBLE, space.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
                 H5Dclose);
        if (!dset.ok()) return fail(SCARL_H5_E_IO, "write_times: cannot create 'times' (exists?)");
        if (H5Dwrite(dset.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, times) < 0)
            return fail(SCARL_H5_E_IO, "write_times: 'times' write failed");
    }
    Hid tid(H5Tcopy(H5T_FORTRAN_S1), H5Tclose);
    if (!tid.ok() || H5Tset_size(tid.id, SCARL_H5_LABEL_LEN) < 0) return fail(SCARL_H5_E_IO, "write_times: string type");
    const hsize_t dims[1] = {(hsize_t)ntimes};
    Hid space(H5Screate_simple(1, dims, nullptr), H5Sclose);
    Hid dset(H5Dcreate2(f->file, "time_labels", tid.id, space.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
    if (!dset.ok()) return fail(SCARL_H5_E_IO, "write_times: cannot create 'time_labels' (exists?)");
    if (H5Dwrite(dset.id, tid.id, space.id, space.id, H5P_DEFAULT, labels) < 0)
        return fail(SCARL_H5_E_IO, "write_times: 'time_labels' write failed");
    return SCARL_H5_OK;
}

int scarl_h5_dataset_info(scarl_h5* f, const char* name, int32_t dims[3], int32_t* file_prec) {
    quiet();
    if (!f || !name || !dims || !file_prec) return fail(SCARL_H5_E_INVALID, "dataset_info: invalid argument");
    Hid dset(H5Dopen2(f->file, name, H5P_DEFAULT), H5Dclose);
    if (!dset.ok()) return fail(SCARL_H5_E_IO, "dataset_info: no dataset %s", name);
    Hid space(H5Dget_space(dset.id), H5Sclose);
    const int nd = H5Sget_simple_extent_ndims(space.id);
    if (nd < 1 || nd > 3) return fail(SCARL_H5_E_INVALID, "dataset_info %s: rank %d", name, nd);
    hsize_t d[3] = {1, 1, 1};
    H5Sget_simple_extent_dims(space.id, d, nullptr);
    for (int a = 0; a < 3; ++a) dims[a] = a < nd ? (int32_t)d[nd - 1 - a] : 1;
    Hid t(H5Dget_type(dset.id), H5Tclose);
    if (H5Tget_class(t.id) != H5T_FLOAT) return fail(SCARL_H5_E_INVALID, "dataset_info %s: not a float dataset", name);
    *file_prec = (int32_t)(8 * H5Tget_size(t.id));
    return SCARL_H5_OK;
}

int scarl_h5_read_block(scarl_h5* f, const char* ds_name, void* data, int32_t mem_prec, int64_t ld1, int64_t ld2,
                        const int32_t offset[3], const int32_t count[3]) {
    quiet();
    if (!f || !ds_name || !data || !offset || !count) return fail(SCARL_H5_E_INVALID, "read_block: invalid argument");
    const hid_t mtype = mem_type(mem_prec);
    if (mtype < 0) return fail(SCARL_H5_E_INVALID, "read_block: precision %d (64 or 32)", mem_prec);
    if (!valid_block(offset, count, ld1, ld2)) return fail(SCARL_H5_E_INVALID, "read_block: block or strides invalid");
    Hid dset(H5Dopen2(f->file, ds_name, H5P_DEFAULT), H5Dclose);
    if (!dset.ok()) return fail(SCARL_H5_E_IO, "read_block: no dataset %s", ds_name);
    hsize_t fd[3];
    std::string why;
    Hid fs(block_filespace(dset.id, offset, count, fd, why), H5Sclose);
    if (!fs.ok()) return fail(SCARL_H5_E_INVALID, "read_block %s: %s", ds_name, why.c_str());
    Hid ms(block_memspace(count, ld1, ld2), H5Sclose);
    if (H5Dread(dset.id, mtype, ms.id, fs.id, H5P_DEFAULT, data) < 0)
        return fail(SCARL_H5_E_IO, "read_block %s: H5Dread failed", ds_name);
    return SCARL_H5_OK;
}

int scarl_h5_read_attr_i32(scarl_h5* f, const char* object, const char* name, int32_t* v, int32_t cap, int32_t* n) {
    return read_attr(f, object, name, H5T_NATIVE_INT32, v, sizeof(int32_t), cap, n);
}

int scarl_h5_read_attr_f64(scarl_h5* f, const char* object, const char* name, double* v, int32_t cap, int32_t* n) {
    return read_attr(f, object, name, H5T_NATIVE_DOUBLE, v, sizeof(double), cap, n);
}

int scarl_h5_attr_count(scarl_h5* f, const char* object, int32_t* n) {
    quiet();
    if (!f || !n) return fail(SCARL_H5_E_INVALID, "attr_count: invalid argument");
    Hid obj(open_object(f, object), H5Oclose);
    if (!obj.ok()) return fail(SCARL_H5_E_IO, "attr_count: cannot open '%s'", object ? object : "");
    H5O_info_t info;
    if (H5Oget_info2(obj.id, &info, H5O_INFO_NUM_ATTRS) < 0) return fail(SCARL_H5_E_IO, "attr_count: info failed");
    *n = (int32_t)info.num_attrs;
    return SCARL_H5_OK;
}

int scarl_h5_attr_name(scarl_h5* f, const char* object, int32_t idx, char* buf, int32_t len) {
    quiet();
    if (!f || !buf || len < 1 || idx < 0) return fail(SCARL_H5_E_INVALID, "attr_name: invalid argument");
    Hid obj(open_object(f, object), H5Oclose);
    if (!obj.ok()) return fail(SCARL_H5_E_IO, "attr_name: cannot open '%s'", object ? object : "");
    const ssize_t r = H5Aget_name_by_idx(obj.id, ".", H5_INDEX_NAME, H5_ITER_INC, (hsize_t)idx, buf, (size_t)len,
                                         H5P_DEFAULT);
    if (r < 0) return fail(SCARL_H5_E_IO, "attr_name: index %d", idx);
    return SCARL_H5_OK;
}

int scarl_h5_read_times(scarl_h5* f, double* times, int32_t cap_times, char* labels, int32_t cap_labels,
                        int32_t* ntimes, int32_t* nprocs) {
    quiet();
    if (!f || !ntimes || !nprocs || cap_times < 0 || cap_labels < 0) return fail(SCARL_H5_E_INVALID, "read_times: invalid argument");
    Hid dset(H5Dopen2(f->file, "times", H5P_DEFAULT), H5Dclose);
    if (!dset.ok()) return fail(SCARL_H5_E_IO, "read_times: no 'times'");
    Hid space(H5Dget_space(dset.id), H5Sclose);
    hsize_t d[2];
    if (H5Sget_simple_extent_ndims(space.id) != 2) return fail(SCARL_H5_E_INVALID, "read_times: 'times' is not rank 2");
    H5Sget_simple_extent_dims(space.id, d, nullptr);
    *nprocs = (int32_t)d[0];
    *ntimes = (int32_t)d[1];
    std::vector<double> t((size_t)(d[0] * d[1]));
    if (H5Dread(dset.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, t.data()) < 0)
        return fail(SCARL_H5_E_IO, "read_times: 'times' read failed");
    if (times) std::memcpy(times, t.data(), sizeof(double) * std::min<size_t>(t.size(), (size_t)cap_times));
    Hid ldset(H5Dopen2(f->file, "time_labels", H5P_DEFAULT), H5Dclose);
    if (!ldset.ok()) return fail(SCARL_H5_E_IO, "read_times: no 'time_labels'");
    Hid lspace(H5Dget_space(ldset.id), H5Sclose);
    const hssize_t nl = H5Sget_simple_extent_npoints(lspace.id);
    Hid tid(H5Tcopy(H5T_FORTRAN_S1), H5Tclose);
    H5Tset_size(tid.id, SCARL_H5_LABEL_LEN);
    std::vector<char> l((size_t)nl * SCARL_H5_LABEL_LEN);
    if (H5Dread(ldset.id, tid.id, H5S_ALL, H5S_ALL, H5P_DEFAULT, l.data()) < 0)
        return fail(SCARL_H5_E_IO, "read_times: 'time_labels' read failed");
    if (labels) std::memcpy(labels, l.data(), std::min<size_t>((size_t)nl, (size_t)cap_labels) * SCARL_H5_LABEL_LEN);
    return SCARL_H5_OK;
}

int scarl_xmf_write_elements(const char* xmf_folder, const char* xmf_name, const char* h5_name,
                             const char* h5_to_xmf_path, const char* ds_name, const double xMin[3],
                             const double xMax[3], const int32_t L[3], int32_t file_prec) {
    if (!xmf_name || !*xmf_name || !h5_name || !ds_name || !xMin || !xMax || !L || (file_prec != 64 && file_prec != 32))
        return fail(SCARL_H5_E_INVALID, "xmf_write_elements: invalid argument");
    // xStep = (xMax - xMin) / dble(xNStep - 1)  (write_output.f90:481)
    double step[3];
    for (int d = 0; d < 3; ++d) step[d] = (xMax[d] - xMin[d]) / (double)(L[d] - 1);
    const std::string dims = dim_text(L), prec = file_prec == 64 ? "8" : "4";
    const std::string h5p = h5_to_xmf_path && *h5_to_xmf_path ? h5_to_xmf_path : ".";
    std::string t;
    t += "<?xml version=\"1.0\" ?>\n";
    t += "<Xdmf Version=\"2.0\">\n";
    t += " <Domain>\n";
    t += "   <DataItem Name=\"RF_1\" Format=\"HDF\" DataType=\"Float\" Precision=\"" + prec + "\" Dimensions=\"" + dims +
         "\">\n";
    t += "        " + h5p + "/" + h5_name + ":/" + ds_name + "\n";
    t += "   </DataItem>\n";
    t += "  <Grid GridType=\"Collection\" CollectionType=\"Spatial\">\n";
    t += "   <Grid Name=\"Group1\">\n";
    t += "     <Topology TopologyType=\"3DCoRectMesh\" Dimensions=\"" + dims + "\"/>\n";
    t += "     <Geometry GeometryType=\"ORIGIN_DXDYDZ\">\n";
    t += "   <DataItem Name=\"origin\" Format=\"XML\" DataType=\"Float\" Precision=\"8\" Dimensions=\"3\">\n";
    f25_records(t, xMin[2], xMin[1], xMin[0]);
    t += "   </DataItem>\n";
    t += "   <DataItem Name=\"step\" Format=\"XML\" DataType=\"Float\" Precision=\"8\" Dimensions=\"3\">\n";
    f25_records(t, step[2], step[1], step[0]);
    t += "   </DataItem>\n";
    t += "     </Geometry>\n";
    t += "     <Attribute Name=\"RF1\" Center=\"Node\" AttributeType=\"Scalar\">\n";
    t += "       <DataItem Reference=\"XML\">\n";
    t += "         /Xdmf/Domain/DataItem[@Name=\"RF_1\"]\n";
    t += "       </DataItem>\n";
    t += "     </Attribute>\n";
    t += "   </Grid>\n";
    t += "  </Grid>\n";
    t += " </Domain>\n";
    t += "</Xdmf>\n";
    return write_text(join_path(xmf_folder, xmf_name), t);
}

int scarl_xmf_write_global(const char* path, const char* const* h5_list, const double* coord_0_list,
                           const double xStep[3], const int32_t xNStep[3], int32_t nb_procs) {
    if (!path || !*path || !h5_list || !coord_0_list || !xStep || !xNStep || nb_procs < 1)
        return fail(SCARL_H5_E_INVALID, "xmf_write_global: invalid argument");
    for (int i = 0; i < nb_procs; ++i)
        if (!h5_list[i]) return fail(SCARL_H5_E_INVALID, "xmf_write_global: h5_list[%d] is null", i);
    const std::string dims = dim_text(xNStep);
    std::string t;
    t += "<?xml version=\"1.0\" ?>\n";
    t += "<Xdmf Version=\"2.0\">\n";
    t += " <Domain>\n";
    for (int i = 0; i < nb_procs; ++i) {
        const std::string p = std::to_string(i);
        t += "   <DataItem Name=\"RF_1_p" + p + "\" Format=\"HDF\" DataType=\"Float\" Precision=\"8\" Dimensions=\"" +
             dims + "\">\n";
        // ':/sample_1_p'//num2str(i-1): num2str's character(len=30) result is not trimmed here
        t += "        ./" + std::string(h5_list[i]) + ":/sample_1_p" + num2str30(i) + "\n";
        t += "   </DataItem>\n";
    }
    t += "  <Grid GridType=\"Collection\" CollectionType=\"Spatial\">\n";
    for (int i = 0; i < nb_procs; ++i) {
        const std::string p = std::to_string(i);
        const double* c = coord_0_list + 3 * (size_t)i;
        t += "   <Grid Name=\"Proc_" + p + "\">\n";
        t += "     <Topology TopologyType=\"3DCoRectMesh\" Dimensions=\"" + dims + "\"/>\n";
        t += "     <Geometry GeometryType=\"ORIGIN_DXDYDZ\">\n";
        t += "   <DataItem Name=\"origin\" Format=\"XML\" DataType=\"Float\" Precision=\"8\" Dimensions=\"3\">\n";
        f25_records(t, c[2], c[1], c[0]);
        t += "   </DataItem>\n";
        t += "   <DataItem Name=\"step\" Format=\"XML\" DataType=\"Float\" Precision=\"8\" Dimensions=\"3\">\n";
        f25_records(t, xStep[2], xStep[1], xStep[0]);
        t += "   </DataItem>\n";
        t += "     </Geometry>\n";
        t += "     <Attribute Name=\"RF_1\" Center=\"Node\" AttributeType=\"Scalar\">\n";
        t += "       <DataItem Reference=\"XML\">\n";
        t += "         /Xdmf/Domain/DataItem[@Name=\"RF_1_p" + p + "\"]\n";
        t += "       </DataItem>\n";
        t += "     </Attribute>\n";
        t += "   </Grid>\n";
    }
    t += "  </Grid>\n";
    t += " </Domain>\n";
    t += "</Xdmf>\n";
    return write_text(path, t);
}

int scarl_fortran_fixed(double v, int32_t w, int32_t d, char* buf, int32_t len) {
    if (!buf || w < 1 || d < 0 || len <= w) return fail(SCARL_H5_E_INVALID, "fortran_fixed: invalid argument");
    const std::string s = fixed(v, w, d);
    std::memcpy(buf, s.c_str(), s.size() + 1);
    return SCARL_H5_OK;
}

}  // extern "C"
