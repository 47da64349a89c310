// HalosShift.h -- drop-in for src/mpi/HalosShift.h: the send / receive
// buffers of the global shift, sized for a full interior (V_int * 4 links).
#pragma once
#include <vector>

#include "../gauge/GaugeField.h"

class HalosShift {
public:
    std::vector<Complex> send;
    std::vector<Complex> recv;
    int L_int;
    size_t V_int;

    explicit HalosShift(const GeometryCB& geo)
        : send(geo.V_int * 4 * 9), recv(geo.V_int * 4 * 9), L_int(geo.L_int), V_int(geo.V_int) {}
};
