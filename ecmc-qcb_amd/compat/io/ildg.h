// ildg.h -- drop-in for src/io/ildg.h:14-20.  The file is read straight into
// HBM by libqcbflow (own LIME parser, device byte swap and transpose) and the
// host GaugeField is filled from there, halos included.
#pragma once
#include <string>

#include "../gauge/GaugeField.h"
#include "../mpi/MpiTopology.h"

std::string generate_ildg_xml(int L_glob);
void save_ildg_clime(const std::string& filename, const std::string& dirpath,
                     const GaugeField& field, const GeometryCB& geo, const mpi::MpiTopology& topo);
void read_ildg_clime(const std::string& filename, const std::string& dirpath, GaugeField& field,
                     const GeometryCB& geo, mpi::MpiTopology& topo);
