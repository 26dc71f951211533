// HDF5 dataset I/O of the dagtm driver, in the reference's layout
// (dagtm.cxx:612-627; h5comm.c:37-209): 2-D matrices stored row-major as
// doubles, scalars as rank-1 datasets of length 1, id/lb as integers.
#pragma once

#include <string>
#include <vector>

#include "dg_init.hpp"

namespace dgh {

class H5File {
   public:
    static H5File open_read(const std::string &path);
    static H5File create(const std::string &path);   // truncates
    static H5File open_append(const std::string &path); // read-write, created if missing
    ~H5File();
    H5File(H5File &&o) noexcept;
    H5File(const H5File &) = delete;
    H5File &operator=(const H5File &) = delete;
    bool ok() const { return id_ >= 0; }

    bool has(const std::string &name) const;
    // any numeric type, converted to double; 1-D datasets read as n x 1
    bool read_matrix(const std::string &name, Mat &m) const;
    bool read_ints(const std::string &name, std::vector<int> &v) const;
    bool read_scalar(const std::string &name, double &x) const;
    bool read_scalar(const std::string &name, int &x) const;
    bool write_matrix(const std::string &name, const Mat &m);
    bool write_vector(const std::string &name, const std::vector<double> &v);
    bool write_ints(const std::string &name, const std::vector<int> &v);
    bool write_scalar(const std::string &name, double x);
    bool write_scalar(const std::string &name, int x);

   private:
    explicit H5File(long long id) : id_(id) {}
    long long id_ = -1;
};

bool have_hdf5();

}  // namespace dgh
