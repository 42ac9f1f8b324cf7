// inc_io.h -- the reference's INC / GSLIB property text formats, parsed and
// printed on the GPU (read_inc_file.cpp, property_writer.cpp, api.cpp:53-251).
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "krige_launch.h"

namespace hpglx {

// Host copy of a file plus the device buffers the text is parsed from.
struct IncText {
    char* host = nullptr;   // pinned, grow-only
    size_t host_cap = 0;
    void release();
};

// Where the data of an INC file begins: the reference's read_prop_name
// (read_inc_file.cpp:8-47) run over the header on the host.  Returns the
// offset of the first byte after the property-name line and the whitespace
// that follows it, or -1 with `err` set.
long long inc_data_offset(const char* text, long long n, std::string& name, std::string& err);

// Reads `filename` into `io.host` (pinned); returns its length or -1.
long long inc_load_file(IncText& io, const char* filename, std::string& err);

// Parses the data section text[0, n) (host pinned memory) into device
// buffers: `size` values, float (as_int = 0: sscanf "%f", mask = value !=
// undef_f) or integer (as_int = 1: sscanf "%d" stored as a byte, mask =
// value != undef_i).  Synchronises `st`; returns 0 or -1 with the
// reference's message in `err`.
int inc_parse(const char* text, long long n, int size, int as_int, float undef_f, int undef_i, float* d_f,
              unsigned char* d_b, unsigned char* d_mask, Workspace& ws, hipStream_t st, std::string& err);

// Remaps informed bytes to the index of their value in `values`
// (api.cpp:94-109); `values` is a host array.
int inc_remap_bytes(unsigned char* d_data, const unsigned char* d_mask, int size, const unsigned char* values,
                    int values_count, Workspace& ws, hipStream_t st, std::string& err);

// Prints `size` values as the reference writes them (property_writer.cpp):
// float -> "%E\n" of the value (undef where the mask is 0); byte ->
// "%d\n" of (signed char)remap[value] (undef where the mask is 0).  The
// text lands in `io.host` (pinned); returns its length or -1.
long long inc_format(const float* d_f, const unsigned char* d_b, const unsigned char* d_mask, int size, float undef_f,
                     unsigned char undef_b, const unsigned char* remap, int remap_count, IncText& io, Workspace& ws,
                     hipStream_t st, std::string& err);

}  // namespace hpglx
