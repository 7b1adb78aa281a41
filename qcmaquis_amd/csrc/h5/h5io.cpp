// Result and checkpoint files (HDF5): the datasets the reference's simulation writes
// (reference: applications/dmrg/dmrg_sim.hpp:88-100,150-160 results; sim/sim.hpp:223-235
// checkpoint status; utils/results_collector.h:56-63 "mean/value" vectors).  Datasets are
// addressed by absolute paths; missing groups are created; writing an existing dataset replaces
// it.  Host code only, linked against the HDF5 C library; separate from the device library.
#include <hdf5.h>

#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "qcmaquis_amd_h5.h"

namespace {

struct Quiet {  // HDF5 prints its error stack by default; errors are return codes here
    H5E_auto2_t fn = nullptr;
    void* data = nullptr;
    Quiet() {
        H5Eget_auto2(H5E_DEFAULT, &fn, &data);
        H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    }
    ~Quiet() { H5Eset_auto2(H5E_DEFAULT, fn, data); }
};

hid_t open_file(const char* file, bool write) {
    if (!write) return H5Fopen(file, H5F_ACC_RDONLY, H5P_DEFAULT);
    hid_t f = H5Fopen(file, H5F_ACC_RDWR, H5P_DEFAULT);
    if (f < 0) f = H5Fcreate(file, H5F_ACC_EXCL, H5P_DEFAULT, H5P_DEFAULT);
    return f;
}

bool exists(hid_t f, const char* path) {
    // every prefix must exist for H5Lexists on the full path
    std::string p(path);
    size_t pos = 1;
    while (pos <= p.size()) {
        size_t next = p.find('/', pos);
        std::string prefix = p.substr(0, next == std::string::npos ? p.size() : next);
        if (prefix.size() > 1 && H5Lexists(f, prefix.c_str(), H5P_DEFAULT) <= 0) return false;
        if (next == std::string::npos) break;
        pos = next + 1;
    }
    return true;
}

// dataset of type t at path with rank 0 (scalar, n == 1) or 1, replacing an existing one
int write(const char* file, const char* path, hid_t mem_type, hid_t file_type, const void* v, uint64_t n, int rank) {
    if (!file || !path || path[0] != '/' || (rank != 0 && rank != 1) || (rank == 0 && n != 1) || (n && !v))
        return MQ_H5_ERR_INVALID;
    Quiet q;
    hid_t f = open_file(file, true);
    if (f < 0) return MQ_H5_ERR_IO;
    int rc = MQ_H5_OK;
    if (exists(f, path) && H5Ldelete(f, path, H5P_DEFAULT) < 0) rc = MQ_H5_ERR_IO;
    hid_t lcpl = H5Pcreate(H5P_LINK_CREATE);
    H5Pset_create_intermediate_group(lcpl, 1);
    hsize_t dims[1] = {hsize_t(n)};
    hid_t space = rank == 0 ? H5Screate(H5S_SCALAR) : H5Screate_simple(1, dims, nullptr);
    hid_t d = rc == MQ_H5_OK ? H5Dcreate2(f, path, file_type, space, lcpl, H5P_DEFAULT, H5P_DEFAULT) : -1;
    if (d < 0) rc = MQ_H5_ERR_IO;
    if (d >= 0 && n && H5Dwrite(d, mem_type, H5S_ALL, H5S_ALL, H5P_DEFAULT, v) < 0) rc = MQ_H5_ERR_IO;
    if (d >= 0) H5Dclose(d);
    H5Sclose(space);
    H5Pclose(lcpl);
    if (H5Fclose(f) < 0) rc = MQ_H5_ERR_IO;
    return rc;
}

struct Dataset {
    hid_t f = -1, d = -1, t = -1, s = -1;
    ~Dataset() {
        if (s >= 0) H5Sclose(s);
        if (t >= 0) H5Tclose(t);
        if (d >= 0) H5Dclose(d);
        if (f >= 0) H5Fclose(f);
    }
    int open(const char* file, const char* path) {
        if (!file || !path) return MQ_H5_ERR_INVALID;
        f = open_file(file, false);
        if (f < 0) return MQ_H5_ERR_IO;
        if (!exists(f, path)) return MQ_H5_ERR_MISSING;
        d = H5Dopen2(f, path, H5P_DEFAULT);
        if (d < 0) return MQ_H5_ERR_MISSING;
        t = H5Dget_type(d);
        s = H5Dget_space(d);
        return MQ_H5_OK;
    }
    uint64_t count() const {
        hssize_t n = H5Sget_simple_extent_npoints(s);
        return n < 0 ? 0 : uint64_t(n);
    }
};

}  // namespace

extern "C" {

int mq_h5_write_f64(const char* file, const char* path, const double* v, uint64_t n, int rank) {
    return write(file, path, H5T_NATIVE_DOUBLE, H5T_IEEE_F64LE, v, n, rank);
}
int mq_h5_write_u64(const char* file, const char* path, const uint64_t* v, uint64_t n, int rank) {
    return write(file, path, H5T_NATIVE_UINT64, H5T_STD_U64LE, v, n, rank);
}
int mq_h5_write_i32(const char* file, const char* path, const int32_t* v, uint64_t n, int rank) {
    return write(file, path, H5T_NATIVE_INT32, H5T_STD_I32LE, v, n, rank);
}

int mq_h5_write_str(const char* file, const char* path, const char* s) {
    if (!s) return MQ_H5_ERR_INVALID;
    Quiet q;
    hid_t t = H5Tcopy(H5T_C_S1);  // variable-length, null-terminated ASCII, as the reference's strings
    H5Tset_size(t, H5T_VARIABLE);
    H5Tset_strpad(t, H5T_STR_NULLTERM);
    H5Tset_cset(t, H5T_CSET_ASCII);
    const char* p = s;
    int rc = write(file, path, t, t, &p, 1, 0);
    H5Tclose(t);
    return rc;
}

int mq_h5_exists(const char* file, const char* path) {
    if (!file || !path) return MQ_H5_ERR_INVALID;
    Quiet q;
    hid_t f = open_file(file, false);
    if (f < 0) return 0;
    int r = exists(f, path) ? 1 : 0;
    H5Fclose(f);
    return r;
}

int mq_h5_info(const char* file, const char* path, int* type_class, int* rank, uint64_t* n) {
    Quiet q;
    Dataset ds;
    int rc = ds.open(file, path);
    if (rc != MQ_H5_OK) return rc;
    H5T_class_t c = H5Tget_class(ds.t);
    if (type_class)
        *type_class = c == H5T_FLOAT ? MQ_H5_FLOAT : c == H5T_INTEGER ? MQ_H5_INTEGER : c == H5T_STRING ? MQ_H5_STRING : 0;
    if (rank) *rank = H5Sget_simple_extent_ndims(ds.s);
    if (n) *n = ds.count();
    return MQ_H5_OK;
}

int mq_h5_read_f64(const char* file, const char* path, double* out, uint64_t cap) {
    Quiet q;
    Dataset ds;
    int rc = ds.open(file, path);
    if (rc != MQ_H5_OK) return rc;
    H5T_class_t c = H5Tget_class(ds.t);
    if (c != H5T_FLOAT && c != H5T_INTEGER) return MQ_H5_ERR_TYPE;
    if (ds.count() > cap || (ds.count() && !out)) return MQ_H5_ERR_INVALID;
    // integer datasets are converted by the library
    if (ds.count() && H5Dread(ds.d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, out) < 0) return MQ_H5_ERR_IO;
    return MQ_H5_OK;
}

int mq_h5_read_i64(const char* file, const char* path, int64_t* out, uint64_t cap) {
    Quiet q;
    Dataset ds;
    int rc = ds.open(file, path);
    if (rc != MQ_H5_OK) return rc;
    if (H5Tget_class(ds.t) != H5T_INTEGER) return MQ_H5_ERR_TYPE;
    if (ds.count() > cap || (ds.count() && !out)) return MQ_H5_ERR_INVALID;
    if (ds.count() && H5Dread(ds.d, H5T_NATIVE_INT64, H5S_ALL, H5S_ALL, H5P_DEFAULT, out) < 0) return MQ_H5_ERR_IO;
    return MQ_H5_OK;
}

int mq_h5_read_str(const char* file, const char* path, char* out, uint64_t cap) {
    Quiet q;
    Dataset ds;
    int rc = ds.open(file, path);
    if (rc != MQ_H5_OK) return rc;
    if (H5Tget_class(ds.t) != H5T_STRING || ds.count() != 1 || !out || cap == 0) return MQ_H5_ERR_TYPE;
    std::string value;
    if (H5Tis_variable_str(ds.t) > 0) {
        hid_t mt = H5Tcopy(H5T_C_S1);
        H5Tset_size(mt, H5T_VARIABLE);
        char* p = nullptr;
        if (H5Dread(ds.d, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, &p) < 0) {
            H5Tclose(mt);
            return MQ_H5_ERR_IO;
        }
        value = p ? p : "";
        H5Dvlen_reclaim(mt, ds.s, H5P_DEFAULT, &p);
        H5Tclose(mt);
    } else {
        size_t sz = H5Tget_size(ds.t);
        std::vector<char> buf(sz + 1, 0);
        if (H5Dread(ds.d, ds.t, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf.data()) < 0) return MQ_H5_ERR_IO;
        value = buf.data();
    }
    if (value.size() + 1 > cap) return MQ_H5_ERR_INVALID;
    std::memcpy(out, value.c_str(), value.size() + 1);
    return MQ_H5_OK;
}

// names of the members of a group, separated by '\n'
int mq_h5_list(const char* file, const char* group, char* out, uint64_t cap) {
    if (!file || !group || !out || cap == 0) return MQ_H5_ERR_INVALID;
    Quiet q;
    hid_t f = open_file(file, false);
    if (f < 0) return MQ_H5_ERR_IO;
    int rc = MQ_H5_OK;
    std::string names;
    hid_t g = (std::strcmp(group, "/") == 0 || exists(f, group)) ? H5Gopen2(f, group, H5P_DEFAULT) : -1;
    if (g < 0) {
        rc = MQ_H5_ERR_MISSING;
    } else {
        H5G_info_t info;
        H5Gget_info(g, &info);
        for (hsize_t i = 0; i < info.nlinks; ++i) {
            ssize_t len = H5Lget_name_by_idx(g, ".", H5_INDEX_NAME, H5_ITER_INC, i, nullptr, 0, H5P_DEFAULT);
            if (len < 0) continue;
            std::vector<char> nm(size_t(len) + 1, 0);
            H5Lget_name_by_idx(g, ".", H5_INDEX_NAME, H5_ITER_INC, i, nm.data(), nm.size(), H5P_DEFAULT);
            if (!names.empty()) names += '\n';
            names += nm.data();
        }
        H5Gclose(g);
    }
    H5Fclose(f);
    if (rc != MQ_H5_OK) return rc;
    if (names.size() + 1 > cap) return MQ_H5_ERR_INVALID;
    std::memcpy(out, names.c_str(), names.size() + 1);
    return MQ_H5_OK;
}

}  // extern "C"
