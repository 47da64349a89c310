#include "ildg.h"

#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <iostream>
#include <vector>

#include "qcbflow.h"

namespace fs = std::filesystem;

std::string generate_ildg_xml(int L_glob) {
    char buf[512];
    std::snprintf(buf, sizeof(buf),
                  "<?xml version=\"1.0\" encoding=\"UTF-8\"?>\n"
                  "<ildgFormat><version>1.0</version><field>su3gauge</field>"
                  "<precision>64</precision><lx>%d</lx><ly>%d</ly><lz>%d</lz><lt>%d</lt>"
                  "</ildgFormat>",
                  L_glob, L_glob, L_glob, L_glob);
    return std::string(buf);
}

namespace {
qcbf_flow* make_handle(const GeometryCB& geo) {
    const int dims[4] = {geo.L_int, geo.L_int, geo.L_int, geo.L_int};
    qcbf_flow* h = nullptr;
    if (qcbf_create(dims, QCBF_FP64, &h) != QCBF_OK) {
        std::cerr << "Fatal: " << qcbf_last_error(nullptr) << std::endl;
        std::exit(1);
    }
    return h;
}
}  // namespace

// save_ildg_clime (src/io/ildg.cpp:34-136): dirpath/filename/filename.
void save_ildg_clime(const std::string& filename, const std::string& dirpath,
                     const GaugeField& field, const GeometryCB& geo, const mpi::MpiTopology&) {
    const fs::path file = fs::path(dirpath) / filename / filename;
    std::error_code ec;
    fs::create_directories(file.parent_path(), ec);
    qcbf_flow* h = make_handle(geo);
    std::vector<double> buf(geo.V_int * 72);
    field.pack_interior(geo, buf.data());
    int rc = qcbf_set_links(h, buf.data());
    if (rc == QCBF_OK) rc = qcbf_save_ildg(h, file.string().c_str());
    if (rc != QCBF_OK) {
        std::cerr << "Fatal: Cannot write " << file << ": " << qcbf_last_error(h) << std::endl;
        qcbf_destroy(h);
        std::exit(1);
    }
    qcbf_destroy(h);
    std::cout << "Configuration saved in " << file << "\n";
}

// read_ildg_clime (src/io/ildg.cpp:138-237): dirpath/filename/filename.
void read_ildg_clime(const std::string& filename, const std::string& dirpath, GaugeField& field,
                     const GeometryCB& geo, mpi::MpiTopology&) {
    const fs::path file = fs::path(dirpath) / filename / filename;
    qcbf_flow* h = make_handle(geo);
    std::vector<double> buf(geo.V_int * 72);
    int rc = qcbf_load_ildg(h, file.string().c_str());
    if (rc == QCBF_OK) rc = qcbf_get_links(h, buf.data());
    if (rc != QCBF_OK) {
        std::cerr << "Erreur : " << file << ": " << qcbf_last_error(h) << std::endl;
        qcbf_destroy(h);
        std::exit(1);
    }
    qcbf_destroy(h);
    field.unpack_interior(geo, buf.data());
    field.fill_halos(geo);
}
