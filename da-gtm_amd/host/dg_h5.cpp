#include "dg_h5.hpp"

#ifdef DG_HAVE_HDF5
#include <hdf5.h>
#endif

namespace dgh {

#ifdef DG_HAVE_HDF5

bool have_hdf5() { return true; }

namespace {
struct Quiet {   // silence HDF5's automatic error stack printing for probes
    H5E_auto2_t f = nullptr;
    void *d = nullptr;
    Quiet() { H5Eget_auto2(H5E_DEFAULT, &f, &d); H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr); }
    ~Quiet() { H5Eset_auto2(H5E_DEFAULT, f, d); }
};

bool write_ds(hid_t f, const std::string &name, hid_t mtype, hid_t ftype, int rank,
              const hsize_t *dims, const void *buf) {
    Quiet q;
    if (H5Lexists(f, name.c_str(), H5P_DEFAULT) > 0) H5Ldelete(f, name.c_str(), H5P_DEFAULT);
    hid_t sp = H5Screate_simple(rank, dims, nullptr);
    if (sp < 0) return false;
    hid_t ds = H5Dcreate2(f, name.c_str(), ftype, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    bool ok = ds >= 0 && H5Dwrite(ds, mtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf) >= 0;
    if (ds >= 0) H5Dclose(ds);
    H5Sclose(sp);
    return ok;
}

bool read_ds(hid_t f, const std::string &name, hid_t mtype, std::vector<hsize_t> &dims,
             std::vector<char> &buf, size_t esize) {
    Quiet q;
    if (H5Lexists(f, name.c_str(), H5P_DEFAULT) <= 0) return false;
    hid_t ds = H5Dopen2(f, name.c_str(), H5P_DEFAULT);
    if (ds < 0) return false;
    hid_t sp = H5Dget_space(ds);
    const int rank = H5Sget_simple_extent_ndims(sp);
    bool ok = rank >= 1 && rank <= 2;
    if (ok) {
        dims.assign(rank, 0);
        H5Sget_simple_extent_dims(sp, dims.data(), nullptr);
        size_t n = 1;
        for (hsize_t d : dims) n *= d;
        buf.resize(n * esize);
        ok = H5Dread(ds, mtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf.data()) >= 0;
    }
    H5Sclose(sp);
    H5Dclose(ds);
    return ok;
}
}  // namespace

H5File H5File::open_read(const std::string &path) {
    Quiet q;
    return H5File(H5Fopen(path.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT));
}

H5File H5File::create(const std::string &path) {
    return H5File(H5Fcreate(path.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT));
}

H5File H5File::open_append(const std::string &path) {
    {
        Quiet q;
        hid_t f = H5Fopen(path.c_str(), H5F_ACC_RDWR, H5P_DEFAULT);
        if (f >= 0) return H5File(f);
    }
    return create(path);
}

H5File::~H5File() {
    if (id_ >= 0) H5Fclose((hid_t)id_);
}

H5File::H5File(H5File &&o) noexcept : id_(o.id_) { o.id_ = -1; }

bool H5File::has(const std::string &name) const {
    Quiet q;
    return id_ >= 0 && H5Lexists((hid_t)id_, name.c_str(), H5P_DEFAULT) > 0;
}

bool H5File::read_matrix(const std::string &name, Mat &m) const {
    std::vector<hsize_t> dims;
    std::vector<char> buf;
    if (!read_ds((hid_t)id_, name, H5T_NATIVE_DOUBLE, dims, buf, sizeof(double))) return false;
    m = Mat((long)dims[0], dims.size() > 1 ? (long)dims[1] : 1);
    std::copy((const double *)buf.data(), (const double *)buf.data() + m.v.size(), m.v.begin());
    return true;
}

bool H5File::read_ints(const std::string &name, std::vector<int> &v) const {
    std::vector<hsize_t> dims;
    std::vector<char> buf;
    if (!read_ds((hid_t)id_, name, H5T_NATIVE_INT, dims, buf, sizeof(int))) return false;
    v.assign((const int *)buf.data(), (const int *)buf.data() + buf.size() / sizeof(int));
    return true;
}

bool H5File::read_scalar(const std::string &name, double &x) const {
    Mat m;
    if (!read_matrix(name, m) || m.v.empty()) return false;
    x = m.v[0];
    return true;
}

bool H5File::read_scalar(const std::string &name, int &x) const {
    std::vector<int> v;
    if (!read_ints(name, v) || v.empty()) return false;
    x = v[0];
    return true;
}

bool H5File::write_matrix(const std::string &name, const Mat &m) {
    hsize_t dims[2] = {(hsize_t)m.rows, (hsize_t)m.cols};
    return write_ds((hid_t)id_, name, H5T_NATIVE_DOUBLE, H5T_IEEE_F64LE, 2, dims, m.data());
}

bool H5File::write_vector(const std::string &name, const std::vector<double> &v) {
    hsize_t dims[1] = {(hsize_t)v.size()};
    return write_ds((hid_t)id_, name, H5T_NATIVE_DOUBLE, H5T_IEEE_F64LE, 1, dims, v.data());
}

bool H5File::write_ints(const std::string &name, const std::vector<int> &v) {
    hsize_t dims[1] = {(hsize_t)v.size()};
    return write_ds((hid_t)id_, name, H5T_NATIVE_INT, H5T_STD_I32LE, 1, dims, v.data());
}

bool H5File::write_scalar(const std::string &name, double x) {
    hsize_t dims[1] = {1};
    return write_ds((hid_t)id_, name, H5T_NATIVE_DOUBLE, H5T_IEEE_F64LE, 1, dims, &x);
}

bool H5File::write_scalar(const std::string &name, int x) {
    hsize_t dims[1] = {1};
    return write_ds((hid_t)id_, name, H5T_NATIVE_INT, H5T_STD_I32LE, 1, dims, &x);
}

#else  // no HDF5 in this build: every call fails cleanly

bool have_hdf5() { return false; }
H5File H5File::open_read(const std::string &) { return H5File(-1); }
H5File H5File::create(const std::string &) { return H5File(-1); }
H5File H5File::open_append(const std::string &) { return H5File(-1); }
H5File::~H5File() {}
H5File::H5File(H5File &&o) noexcept : id_(o.id_) { o.id_ = -1; }
bool H5File::has(const std::string &) const { return false; }
bool H5File::read_matrix(const std::string &, Mat &) const { return false; }
bool H5File::read_ints(const std::string &, std::vector<int> &) const { return false; }
bool H5File::read_scalar(const std::string &, double &) const { return false; }
bool H5File::read_scalar(const std::string &, int &) const { return false; }
bool H5File::write_matrix(const std::string &, const Mat &) { return false; }
bool H5File::write_vector(const std::string &, const std::vector<double> &) { return false; }
bool H5File::write_ints(const std::string &, const std::vector<int> &) { return false; }
bool H5File::write_scalar(const std::string &, double) { return false; }
bool H5File::write_scalar(const std::string &, int) { return false; }

#endif

}  // namespace dgh
