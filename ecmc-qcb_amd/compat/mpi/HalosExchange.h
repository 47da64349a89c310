// HalosExchange.h -- drop-in for src/mpi/HalosExchange.h on one rank: the
// cascade exchange (HalosExchange.cpp:4-72) becomes a periodic copy of the
// opposite interior faces into the halos, edges and corners included.
#pragma once
#include "../gauge/GaugeField.h"
#include "MpiTopology.h"

namespace mpi::exchange {
inline void exchange_halos_cascade(GaugeField& field, const GeometryCB& geo, MpiTopology&) {
    field.fill_halos(geo);
}
}  // namespace mpi::exchange
