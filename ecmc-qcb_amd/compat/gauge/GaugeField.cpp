#include "GaugeField.h"

#include <cstring>

GaugeField::GaugeField(const GeometryCB& geo)
    : L_ext(geo.L_ext), L_int(geo.L_int), V_ext(geo.V_ext), V_int(geo.V_int),
      links(geo.V_ext * 4 * 9, Complex(0.0, 0.0)) {
    for (size_t s = 0; s < V_ext; s++)
        for (int mu = 0; mu < 4; mu++)
            for (int i = 0; i < 3; i++) links[(s * 4 + mu) * 9 + 4 * i] = Complex(1.0, 0.0);
}

void GaugeField::pack_interior(const GeometryCB& geo, double* out) const {
    size_t k = 0;
    for (int t = 1; t <= L_int; t++)
        for (int z = 1; z <= L_int; z++)
            for (int y = 1; y <= L_int; y++) {
                // one x-row of the interior is contiguous in both layouts
                const size_t s = geo.index(1, y, z, t);
                std::memcpy(out + k, &links[s * 36], sizeof(Complex) * 36 * L_int);
                k += 72 * static_cast<size_t>(L_int);
            }
}

void GaugeField::unpack_interior(const GeometryCB& geo, const double* in) {
    size_t k = 0;
    for (int t = 1; t <= L_int; t++)
        for (int z = 1; z <= L_int; z++)
            for (int y = 1; y <= L_int; y++) {
                const size_t s = geo.index(1, y, z, t);
                std::memcpy(static_cast<void*>(&links[s * 36]), in + k, sizeof(Complex) * 36 * L_int);
                k += 72 * static_cast<size_t>(L_int);
            }
}

void GaugeField::fill_halos(const GeometryCB& geo) {
    auto wrap = [&](int c) { return c == 0 ? L_int : (c == L_ext - 1 ? 1 : c); };
    for (int t = 0; t < L_ext; t++)
        for (int z = 0; z < L_ext; z++)
            for (int y = 0; y < L_ext; y++)
                for (int x = 0; x < L_ext; x++) {
                    const int wx = wrap(x), wy = wrap(y), wz = wrap(z), wt = wrap(t);
                    if (wx == x && wy == y && wz == z && wt == t) continue;
                    const size_t dst = geo.index(x, y, z, t), src = geo.index(wx, wy, wz, wt);
                    std::memcpy(static_cast<void*>(&links[dst * 36]), &links[src * 36], sizeof(Complex) * 36);
                }
}
