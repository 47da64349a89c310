// ildg.hpp -- LIME record framing and the ILDG format record (host side).
//
// Written from the public LIME specification (the reference links the c-lime
// library for this: src/io/ildg.cpp:7-9,54-73,155-174).  A LIME record is a
// 144-byte big-endian header
//   bytes 0-3   magic 0x456789ab
//   bytes 4-5   format version (1)
//   bytes 6-7   flags: 0x8000 message-begin (MB), 0x4000 message-end (ME)
//   bytes 8-15  payload length in bytes
//   bytes 16-143 record type, NUL-padded ASCII
// followed by the payload, padded with zeros to a multiple of 8 bytes.
#pragma once
#include <stdint.h>
#include <stdio.h>

#include <string>

namespace qcb {

constexpr uint32_t kLimeMagic = 0x456789abu;
constexpr int kLimeHeaderBytes = 144;

struct ildg_info {
    long long data_offset = -1;    // file offset of the ildg-binary-data payload
    unsigned long long data_bytes = 0;
    bool has_format = false;       // an ildg-format record was found
    int dims[4] = {0, 0, 0, 0};    // lx, ly, lz, lt from ildg-format
    int precision = 64;            // 32 or 64 from ildg-format
    std::string field;             // <field> of ildg-format
};

// Scan every record of an open LIME file.  Returns 0 or a qcbf_status code
// with a message in *err.
int ildg_scan(FILE* fp, ildg_info* info, std::string* err);

// The ildg-format XML the reference writes (generate_ildg_xml,
// src/io/ildg.cpp:13-22), generalised to anisotropic extents.
std::string ildg_format_xml(const int dims[4], int precision);

// Write one LIME record header.  Returns false on a short write.
bool lime_write_header(FILE* fp, const char* type, uint64_t len, bool mb, bool me);

}  // namespace qcb
