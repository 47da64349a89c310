// Shift.h -- drop-in for src/mpi/Shift.h on one rank.
//
// The global shift U_mu(x) <- U_mu(x - s) of a host GaugeField, with the
// reference's entry points and semantics: a shift of s sites along dim moves
// the last s interior slices of that dim to the front (pack, "send" to the
// next rank, which on one rank is this one, unpack) and the others up by s
// (apply_local_shift).  random_shift draws the four shifts with the caller's
// std::mt19937_64 exactly as Shift.cpp:163-186 does, prints the same line and
// refreshes the halos.  A field that lives on the device is shifted there by
// qcbf_shift (include/qcbflow.h).
#pragma once
#include <random>
#include <vector>

#include "../gauge/GaugeField.h"
#include "HalosShift.h"
#include "MpiTopology.h"

namespace mpi::shift {
// Copies the interior slices start..end (1-based) of dim into buffer, in the
// reference's order (t, z, y, x loops, x fastest; 4 links of 9 entries).
void pack_shift_slice(const GaugeField& field, const GeometryCB& geo, std::vector<Complex>& buffer,
                      int dim, int start, int end);
void unpack_shift_slice(GaugeField& field, const GeometryCB& geo,
                        const std::vector<Complex>& buffer, int dim, int start, int end);
// Moves interior slice c to c + s along dim for c = 1 .. L - s (no wrap).
void apply_local_shift(GaugeField& field, const GeometryCB& geo, int dim, int s);
// 0 <= s[d] < L for every d.
void shift_field(GaugeField& field, const GeometryCB& geo, HalosShift& h, mpi::MpiTopology& topo,
                 const int s[4]);
void random_shift(GaugeField& field, const GeometryCB& geo, HalosShift& h, mpi::MpiTopology& topo,
                  std::mt19937_64& rng);
}  // namespace mpi::shift
