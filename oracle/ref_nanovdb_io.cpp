// ref_nanovdb_io.cpp — test-only shim over the reference's own NanoVDB file IO
// (extern/nanovdb/util/IO.h, as src/App/loaders/load_volumes.cpp:100-109 uses
// it through nanovdb::io::readGrid), so the tests can check the .nvdb reader
// and writer of stratum2_amd/assets.py against it byte for byte.  Built by
// oracle/Makefile into oracle/_ref/ only when the reference tree is present.
#include <cstring>
#include <stdexcept>

#include "util/IO.h"

extern "C" {

// Reads grid n of a .nvdb file into out (capacity cap bytes); returns the grid
// size in bytes, 0 on failure.
unsigned long long ref_nvdb_read(const char* path, unsigned long long n, void* out, unsigned long long cap) {
    try {
        auto h = nanovdb::io::readGrid(path, n);
        if (h.size() > cap) return 0;
        std::memcpy(out, h.data(), h.size());
        return h.size();
    } catch (const std::exception&) {
        return 0;
    }
}

// Writes one float grid (bytes) to a .nvdb file; codec 0 none, 1 zip.  Returns 0 on success.
int ref_nvdb_write(const char* path, const void* grid, unsigned long long size, int codec) {
    try {
        auto buf = nanovdb::HostBuffer::create(size);
        std::memcpy(buf.data(), grid, size);
        nanovdb::GridHandle<nanovdb::HostBuffer> h(std::move(buf));
        nanovdb::io::writeGrid(path, h, codec == 1 ? nanovdb::io::Codec::ZIP : nanovdb::io::Codec::NONE);
        return 0;
    } catch (const std::exception&) {
        return 1;
    }
}
}
