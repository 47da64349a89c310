// Host build of the exponential's Cayley-Hamilton coefficients (su3.hpp):
// reads (t, d) pairs as raw doubles from stdin and writes, per pair, the
// coefficients q0, q1, q2 (complex, six doubles each) of horner_n<10>,
// horner_n<18>, horner_exp, horner_n<7> and horner_exp for an FP32 result.
// A checker for tests/test_horner_short.py only.
#include <cstdio>

#include "su3.hpp"

int main() {
    double td[2];
    while (std::fread(td, sizeof(double), 2, stdin) == 2) {
        qcb::cpx<double> q[15];
        qcb::horner_n<10>(td[0], td[1], q[0], q[1], q[2]);
        qcb::horner_n<18>(td[0], td[1], q[3], q[4], q[5]);
        qcb::horner_exp(td[0], td[1], q[6], q[7], q[8]);
        qcb::horner_n<7>(td[0], td[1], q[9], q[10], q[11]);
        qcb::horner_exp<double, true>(td[0], td[1], q[12], q[13], q[14]);
        std::fwrite(q, sizeof(q), 1, stdout);
    }
    return 0;
}
