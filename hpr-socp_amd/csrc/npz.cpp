// .npz writer (ZIP, method "stored", ZIP64 fields on every member).  Layout:
//   [local header + name + zip64 extra + .npy header + data] per member
//   [central directory entry + name + zip64 extra] per member
//   [zip64 end-of-central-directory record][zip64 locator][end-of-central-directory]
// following the PKWARE APPNOTE (sections 4.3.7, 4.3.12, 4.3.14-4.3.16, 4.5.3)
// and the .npy format description of numpy (version 1.0 headers).
#include "npz.h"

#include <cstdio>
#include <cstring>
#include <ctime>

namespace hprsocp {

namespace {

// Slicing-by-8 tables of the reflected CRC-32 (Kounavis and Berry, "A Systematic
// Approach to Building High Performance Software-based CRC Generators", 2005).
struct CrcTables {
    uint32_t t[8][256];
    CrcTables() {
        for (uint32_t i = 0; i < 256; ++i) {
            uint32_t c = i;
            for (int k = 0; k < 8; ++k) c = (c & 1) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
            t[0][i] = c;
        }
        for (uint32_t i = 0; i < 256; ++i)
            for (int s = 1; s < 8; ++s) t[s][i] = (t[s - 1][i] >> 8) ^ t[0][t[s - 1][i] & 0xFF];
    }
};
const CrcTables &crc_tables() {
    static const CrcTables T;
    return T;
}

struct Buf {
    std::vector<uint8_t> b;
    void u16(uint32_t v) {
        b.push_back((uint8_t)v);
        b.push_back((uint8_t)(v >> 8));
    }
    void u32(uint32_t v) {
        for (int k = 0; k < 4; ++k) b.push_back((uint8_t)(v >> (8 * k)));
    }
    void u64(uint64_t v) {
        for (int k = 0; k < 8; ++k) b.push_back((uint8_t)(v >> (8 * k)));
    }
    void bytes(const void *p, size_t n) { b.insert(b.end(), (const uint8_t *)p, (const uint8_t *)p + n); }
};

// MS-DOS date and time of "now" (ZIP's modification stamp)
void dos_stamp(uint16_t *dtime, uint16_t *ddate) {
    const time_t t = time(nullptr);
    struct tm tm_;
    localtime_r(&t, &tm_);
    const int year = tm_.tm_year + 1900 < 1980 ? 1980 : tm_.tm_year + 1900;
    *dtime = (uint16_t)((tm_.tm_hour << 11) | (tm_.tm_min << 5) | (tm_.tm_sec / 2));
    *ddate = (uint16_t)(((year - 1980) << 9) | ((tm_.tm_mon + 1) << 5) | tm_.tm_mday);
}

const uint32_t kMax32 = 0xFFFFFFFFu;

}  // namespace

uint32_t crc32_update(uint32_t crc, const void *data, size_t len) {
    const CrcTables &T = crc_tables();
    const uint8_t *p = (const uint8_t *)data;
    uint32_t c = ~crc;
    while (len >= 8) {
        uint32_t lo, hi;
        memcpy(&lo, p, 4);
        memcpy(&hi, p + 4, 4);
        lo ^= c;
        c = T.t[7][lo & 0xFF] ^ T.t[6][(lo >> 8) & 0xFF] ^ T.t[5][(lo >> 16) & 0xFF] ^ T.t[4][lo >> 24] ^
            T.t[3][hi & 0xFF] ^ T.t[2][(hi >> 8) & 0xFF] ^ T.t[1][(hi >> 16) & 0xFF] ^ T.t[0][hi >> 24];
        p += 8;
        len -= 8;
    }
    while (len--) c = T.t[0][(c ^ *p++) & 0xFF] ^ (c >> 8);
    return ~c;
}

void NpzWriter::add(const std::string &name, const char *descr, const std::string &shape, const void *data,
                    int64_t bytes, bool copy) {
    Member m;
    m.name = name;
    std::string dict = std::string("{'descr': '") + descr + "', 'fortran_order': False, 'shape': " + shape + ", }";
    // magic (6) + version (2) + header length (2) + dict + padding + '\n', a multiple of 64
    size_t total = 10 + dict.size() + 1;
    dict.append((64 - total % 64) % 64, ' ');
    dict.push_back('\n');
    const uint16_t hl = (uint16_t)dict.size();
    m.header.assign("\x93NUMPY\x01\x00", 8);
    m.header.push_back((char)(hl & 0xFF));
    m.header.push_back((char)(hl >> 8));
    m.header += dict;
    m.bytes = bytes;
    if (copy) m.own.assign((const uint8_t *)data, (const uint8_t *)data + bytes);
    else m.data = data;
    members_.push_back(std::move(m));
}

void NpzWriter::add_f64(const std::string &name, const double *v, int64_t n) {
    add(name, "<f8", "(" + std::to_string(n) + ",)", v, n * (int64_t)sizeof(double), false);
}
void NpzWriter::add_f64(const std::string &name, double v) { add(name, "<f8", "()", &v, sizeof v, true); }
void NpzWriter::add_i64(const std::string &name, int64_t v) { add(name, "<i8", "()", &v, sizeof v, true); }
void NpzWriter::add_i32(const std::string &name, int32_t v) { add(name, "<i4", "()", &v, sizeof v, true); }
void NpzWriter::add_bool(const std::string &name, bool v) {
    const uint8_t b = v ? 1 : 0;
    add(name, "|b1", "()", &b, 1, true);
}
void NpzWriter::add_str(const std::string &name, const std::string &v) {
    // numpy's unicode dtype: UTF-32 little endian code units, no terminator
    std::vector<uint8_t> u(v.size() * 4, 0);
    for (size_t k = 0; k < v.size(); ++k) u[4 * k] = (uint8_t)v[k];
    const std::string descr = "<U" + std::to_string(v.empty() ? 1 : v.size());
    if (v.empty()) u.assign(4, 0);
    add(name, descr.c_str(), "()", u.data(), (int64_t)u.size(), true);
}

bool NpzWriter::write(const std::string &path, std::string *err) const {
    const std::string tmp = path + ".tmp";
    FILE *f = fopen(tmp.c_str(), "wb");
    if (!f) {
        if (err) *err = "cannot open " + tmp + " for writing";
        return false;
    }
    uint16_t dtime, ddate;
    dos_stamp(&dtime, &ddate);
    struct Entry {
        std::string fname;
        uint32_t crc;
        uint64_t size, offset;
    };
    std::vector<Entry> cd;
    uint64_t pos = 0;
    bool ok = true;
    auto put = [&](const void *p, size_t n) {
        if (ok && n > 0 && fwrite(p, 1, n, f) != n) ok = false;
        pos += n;
    };
    for (const Member &m : members_) {
        const uint8_t *payload = m.own.empty() ? (const uint8_t *)m.data : m.own.data();
        Entry e;
        e.fname = m.name + ".npy";
        e.size = m.header.size() + (uint64_t)m.bytes;
        e.offset = pos;
        e.crc = crc32_update(crc32_update(0, m.header.data(), m.header.size()), payload, (size_t)m.bytes);
        Buf h;
        h.u32(0x04034b50);
        h.u16(45);  // version needed: ZIP64
        h.u16(0);   // flags
        h.u16(0);   // method: stored
        h.u16(dtime);
        h.u16(ddate);
        h.u32(e.crc);
        h.u32(kMax32);  // sizes in the ZIP64 extra field
        h.u32(kMax32);
        h.u16((uint32_t)e.fname.size());
        h.u16(20);
        h.bytes(e.fname.data(), e.fname.size());
        h.u16(0x0001);
        h.u16(16);
        h.u64(e.size);
        h.u64(e.size);
        put(h.b.data(), h.b.size());
        put(m.header.data(), m.header.size());
        put(payload, (size_t)m.bytes);
        cd.push_back(e);
    }
    const uint64_t cd_offset = pos;
    Buf c;
    for (const Entry &e : cd) {
        c.u32(0x02014b50);
        c.u16(45 | (3 << 8));  // made by: UNIX, ZIP64
        c.u16(45);
        c.u16(0);
        c.u16(0);
        c.u16(dtime);
        c.u16(ddate);
        c.u32(e.crc);
        c.u32(kMax32);
        c.u32(kMax32);
        c.u16((uint32_t)e.fname.size());
        c.u16(28);
        c.u16(0);  // comment
        c.u16(0);  // disk
        c.u16(0);  // internal attributes
        c.u32(0100644u << 16);
        c.u32(kMax32);  // local header offset in the extra field
        c.bytes(e.fname.data(), e.fname.size());
        c.u16(0x0001);
        c.u16(24);
        c.u64(e.size);
        c.u64(e.size);
        c.u64(e.offset);
    }
    const uint64_t cd_size = c.b.size();
    const uint64_t eocd64 = cd_offset + cd_size;
    const uint64_t count = cd.size();
    c.u32(0x06064b50);
    c.u64(44);
    c.u16(45 | (3 << 8));
    c.u16(45);
    c.u32(0);
    c.u32(0);
    c.u64(count);
    c.u64(count);
    c.u64(cd_size);
    c.u64(cd_offset);
    c.u32(0x07064b50);
    c.u32(0);
    c.u64(eocd64);
    c.u32(1);
    c.u32(0x06054b50);
    c.u16(0);
    c.u16(0);
    c.u16(count >= 0xFFFF ? 0xFFFF : (uint32_t)count);
    c.u16(count >= 0xFFFF ? 0xFFFF : (uint32_t)count);
    c.u32(kMax32);
    c.u32(kMax32);
    c.u16(0);
    put(c.b.data(), c.b.size());
    if (fclose(f) != 0) ok = false;
    if (!ok) {
        remove(tmp.c_str());
        if (err) *err = "write to " + tmp + " failed";
        return false;
    }
    if (rename(tmp.c_str(), path.c_str()) != 0) {
        remove(tmp.c_str());
        if (err) *err = "cannot rename " + tmp + " to " + path;
        return false;
    }
    return true;
}

}  // namespace hprsocp
