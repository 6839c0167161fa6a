// File helpers shared by the on-disk ELL format and the Lanczos checkpoint.
#pragma once

#include <cstdio>
#include <string>

#include "internal.hpp"

namespace qmb {

constexpr size_t kStageBytes = (size_t)64 << 20;  // pinned staging per copy

struct File {
    std::FILE *f;
    std::string name;
    File(const char *path, const char *mode);
    ~File();
    File(const File &) = delete;
    File &operator=(const File &) = delete;
    void write(const void *p, size_t n);
    void read(void *p, size_t n);
    void close();  // reports a failed flush
};

// Device <-> file through a pinned staging buffer (synchronous copies).
void write_device(File &out, const void *dev, size_t bytes);
void read_device(File &in, void *dev, size_t bytes);

}  // namespace qmb
