#include "io.h"

#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>

namespace fs = std::filesystem;

std::string io::format_double(double val, int precision) {
    std::stringstream ss;
    ss << std::fixed << std::setprecision(precision) << val;
    return ss.str();
}

void io::save_topo(const std::vector<double>& tQE, const std::string& filename,
                   const std::string& dirpath, int precision) {
    const fs::path dir = fs::path(dirpath) / filename;
    std::error_code ec;
    if (!fs::exists(dir)) fs::create_directories(dir, ec);
    if (ec) {
        std::cerr << "Couldn't create folder data : " << ec.message() << std::endl;
        return;
    }
    const fs::path filepath = dir / (filename + "_topo.txt");
    std::ofstream file(filepath, std::ios::out | std::ios::app);
    if (!file.is_open()) {
        std::cout << "Could not open file " << filepath << "\n";
        return;
    }
    file << std::fixed << std::setprecision(precision);
    for (size_t i = 0; i + 2 < tQE.size(); i += 3)
        file << tQE[i] << " " << tQE[i + 1] << " " << tQE[i + 2] << "\n";
    std::cout << "Topology written in " << filepath << "\n";
}
