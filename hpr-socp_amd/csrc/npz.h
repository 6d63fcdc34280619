// Writer of uncompressed .npz archives: a ZIP file whose members are .npy
// arrays, readable by numpy.load with allow_pickle=False.  The solver's
// auto-save checkpoint (the current / best / parameters groups of the
// reference's HDF5 file, reference src/algorithm.jl:352-418) is written with
// it; member names keep the reference's dataset paths ("best/x_org").
//
// Every member carries ZIP64 size and offset fields, so an archive with
// vectors beyond 4 GiB takes the same code path as a small one.  The archive
// is written to `path + ".tmp"` and renamed over `path`, so a reader never
// sees a partly written checkpoint.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace hprsocp {

class NpzWriter {
public:
    // shape (n,); the data must stay valid until write() returns
    void add_f64(const std::string &name, const double *v, int64_t n);
    // shape () scalars
    void add_f64(const std::string &name, double v);
    void add_i64(const std::string &name, int64_t v);
    void add_i32(const std::string &name, int32_t v);
    void add_bool(const std::string &name, bool v);
    // a '<U' string scalar; the text must be ASCII
    void add_str(const std::string &name, const std::string &v);
    bool write(const std::string &path, std::string *err) const;

private:
    struct Member {
        std::string name;    // without the ".npy" suffix
        std::string header;  // complete .npy header
        const void *data = nullptr;
        int64_t bytes = 0;
        std::vector<uint8_t> own;  // payload of scalars and strings
    };
    void add(const std::string &name, const char *descr, const std::string &shape, const void *data, int64_t bytes,
             bool copy);
    std::vector<Member> members_;
};

// CRC-32 of the ZIP format (reflected polynomial 0xEDB88320), continued from crc.
uint32_t crc32_update(uint32_t crc, const void *data, size_t len);

}  // namespace hprsocp
