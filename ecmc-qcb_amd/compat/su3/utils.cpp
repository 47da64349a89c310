#include "utils.h"

#include <iostream>

#include "../flow/gradient_flow.h"

SU3 exp_su3_luscher(const SU3& Z, double coeff) {
    double in[18], out[18];
    for (int k = 0; k < 9; k++) {
        in[2 * k] = Z.data()[k].real();
        in[2 * k + 1] = Z.data()[k].imag();
    }
    int flagged = 0;
    qcbf_compat_check(qcbf_su3_exp(in, 1, coeff, out, &flagged), nullptr, "qcbf_su3_exp");
    if (flagged) std::cerr << "Warning : ||X|| > 1, unstable exp\n";  // utils.cpp:153-155
    SU3 r;
    for (int k = 0; k < 9; k++) r.data()[k] = Complex(out[2 * k], out[2 * k + 1]);
    return r;
}
