// Test driver for the library's .npz writer (hpr-socp_amd/csrc/npz.cpp):
// writes one archive with every member type the auto-save checkpoint uses,
// plus a CRC-32 check value, so test_npz.py can read it back with numpy.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "npz.h"

int main(int argc, char **argv) {
    if (argc < 3) return 2;
    const std::string path = argv[1];
    const long n = atol(argv[2]);
    std::vector<double> v((size_t)n);
    for (long i = 0; i < n; ++i) v[(size_t)i] = 0.5 * (double)i - 3.0;
    hprsocp::NpzWriter w;
    w.add_f64("current/x_org", v.data(), n);
    w.add_f64("current/sigma", 0.125);
    w.add_i64("current/iteration", 123456789012LL);
    w.add_i32("parameters/device_number", -7);
    w.add_bool("parameters/auto_save", true);
    w.add_bool("parameters/warm_up", false);
    w.add_str("parameters/bc_scaling_norm_type", "l2");
    w.add_str("current/timestamp", "2026-12-03T10:11:12.345");
    w.add_f64("best/y_org", v.data(), 0);
    std::string err;
    if (!w.write(path, &err)) {
        fprintf(stderr, "%s\n", err.c_str());
        return 1;
    }
    // CRC-32 check value of the ZIP polynomial: "123456789" -> 0xCBF43926
    printf("%08x\n", hprsocp::crc32_update(0, "123456789", 9));
    // continued CRC equals the one-shot CRC
    const char *s = "The quick brown fox jumps over the lazy dog";
    const uint32_t a = hprsocp::crc32_update(0, s, strlen(s));
    const uint32_t b = hprsocp::crc32_update(hprsocp::crc32_update(0, s, 13), s + 13, strlen(s) - 13);
    printf("%08x %08x\n", a, b);
    return 0;
}
