// ildg.cpp -- LIME record scanner / writer and the ildg-format record.
#include "ildg.hpp"

#include <string.h>

#include <vector>

#include "qcbflow.h"

namespace qcb {

static uint64_t be64(const unsigned char* p) {
    uint64_t v = 0;
    for (int k = 0; k < 8; k++) v = (v << 8) | p[k];
    return v;
}
static uint32_t be32(const unsigned char* p) {
    return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3];
}

static bool xml_int(const std::string& xml, const char* tag, int* out) {
    const std::string open = std::string("<") + tag + ">";
    const size_t a = xml.find(open);
    if (a == std::string::npos) return false;
    const size_t b = xml.find("</", a + open.size());
    if (b == std::string::npos) return false;
    const std::string v = xml.substr(a + open.size(), b - a - open.size());
    char* end = nullptr;
    const long x = strtol(v.c_str(), &end, 10);
    if (end == v.c_str()) return false;
    *out = (int)x;
    return true;
}

int ildg_scan(FILE* fp, ildg_info* info, std::string* err) {
    if (fseeko(fp, 0, SEEK_END) != 0) {
        *err = "cannot seek in file";
        return QCBF_ERR_IO;
    }
    const long long fsize = ftello(fp);
    long long pos = 0;
    int nrec = 0;
    while (pos + kLimeHeaderBytes <= fsize) {
        unsigned char h[kLimeHeaderBytes];
        if (fseeko(fp, pos, SEEK_SET) != 0 || fread(h, 1, kLimeHeaderBytes, fp) != kLimeHeaderBytes) {
            *err = "short read of a LIME record header";
            return QCBF_ERR_IO;
        }
        if (be32(h) != kLimeMagic) {
            *err = "bad LIME magic number at offset " + std::to_string(pos);
            return QCBF_ERR_FORMAT;
        }
        const uint64_t len = be64(h + 8);
        char type[129];
        memcpy(type, h + 16, 128);
        type[128] = '\0';
        const long long data = pos + kLimeHeaderBytes;
        if ((unsigned long long)(fsize - data) < len) {
            *err = std::string("LIME record '") + type + "' is truncated";
            return QCBF_ERR_FORMAT;
        }
        nrec++;
        if (strcmp(type, "ildg-format") == 0) {
            if (len > (1u << 20)) {
                *err = "ildg-format record too large";
                return QCBF_ERR_FORMAT;
            }
            std::string xml(len, '\0');
            if (len && fread(&xml[0], 1, len, fp) != len) {
                *err = "short read of ildg-format record";
                return QCBF_ERR_IO;
            }
            info->has_format = true;
            int v = 0;
            const char* tags[4] = {"lx", "ly", "lz", "lt"};
            for (int d = 0; d < 4; d++) {
                if (!xml_int(xml, tags[d], &v)) {
                    *err = std::string("ildg-format record lacks <") + tags[d] + ">";
                    return QCBF_ERR_FORMAT;
                }
                info->dims[d] = v;
            }
            if (xml_int(xml, "precision", &v)) info->precision = v;
            const size_t a = xml.find("<field>"), b = xml.find("</field>");
            if (a != std::string::npos && b != std::string::npos && b > a)
                info->field = xml.substr(a + 7, b - a - 7);
        } else if (strcmp(type, "ildg-binary-data") == 0 && info->data_offset < 0) {
            info->data_offset = data;
            info->data_bytes = len;
        }
        pos = data + (long long)((len + 7) & ~7ull);
    }
    if (nrec == 0) {
        *err = "no LIME record in file";
        return QCBF_ERR_FORMAT;
    }
    if (info->data_offset < 0) {
        *err = "record 'ildg-binary-data' not found";
        return QCBF_ERR_FORMAT;
    }
    return QCBF_OK;
}

std::string ildg_format_xml(const int dims[4], int precision) {
    char buf[512];
    snprintf(buf, sizeof(buf),
             "<?xml version=\"1.0\" encoding=\"UTF-8\"?>\n"
             "<ildgFormat><version>1.0</version><field>su3gauge</field>"
             "<precision>%d</precision><lx>%d</lx><ly>%d</ly><lz>%d</lz><lt>%d</lt>"
             "</ildgFormat>",
             precision, dims[0], dims[1], dims[2], dims[3]);
    return std::string(buf);
}

bool lime_write_header(FILE* fp, const char* type, uint64_t len, bool mb, bool me) {
    unsigned char h[kLimeHeaderBytes];
    memset(h, 0, sizeof(h));
    h[0] = 0x45;
    h[1] = 0x67;
    h[2] = 0x89;
    h[3] = 0xab;
    h[4] = 0x00;
    h[5] = 0x01;
    h[6] = (unsigned char)((mb ? 0x80 : 0) | (me ? 0x40 : 0));
    h[7] = 0x00;
    for (int k = 0; k < 8; k++) h[8 + k] = (unsigned char)(len >> (56 - 8 * k));
    strncpy((char*)h + 16, type, 127);
    return fwrite(h, 1, sizeof(h), fp) == sizeof(h);
}

}  // namespace qcb
