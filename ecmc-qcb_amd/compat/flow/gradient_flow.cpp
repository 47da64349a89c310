#include "gradient_flow.h"

#include <cstdio>
#include <cstdlib>

void qcbf_compat_check(int rc, const qcbf_flow* h, const char* what) {
    if (rc == QCBF_OK) return;
    std::fprintf(stderr, "Fatal: %s failed (%d): %s\n", what, rc, qcbf_last_error(h));
    std::exit(1);
}

GradientFlow::GradientFlow(double epsilon_, const GaugeField& field, const GeometryCB& geo)
    : epsilon(epsilon_), field_c(field), force_0(field), force_1(field), geo_p(&geo),
      buf_(geo.V_int * 72) {
    const int dims[4] = {geo.L_int, geo.L_int, geo.L_int, geo.L_int};
    qcbf_compat_check(qcbf_create(dims, QCBF_FP64, &h_), nullptr, "qcbf_create");
    field.pack_interior(geo, buf_.data());
    qcbf_compat_check(qcbf_set_links(h_, buf_.data()), h_, "qcbf_set_links");
}

GradientFlow::~GradientFlow() { qcbf_destroy(h_); }

// GradientFlow::copy (src/flow/gradient_flow.cpp:138): field_c = field.
void GradientFlow::copy(const GaugeField& field) {
    field.pack_interior(*geo_p, buf_.data());
    qcbf_compat_check(qcbf_set_links(h_, buf_.data()), h_, "qcbf_set_links");
    if (mirror_host) field_c = field;
}

void GradientFlow::download_field_c() {
    qcbf_compat_check(qcbf_get_links(h_, buf_.data()), h_, "qcbf_get_links");
    field_c.unpack_interior(*geo_p, buf_.data());
}

void GradientFlow::download_force(int i) {
    qcbf_compat_check(qcbf_get_force(h_, i, buf_.data()), h_, "qcbf_get_force");
    (i == 0 ? force_0 : force_1).unpack_interior(*geo_p, buf_.data());
}

void GradientFlow::compute_force(int i) {
    qcbf_compat_check(qcbf_compute_force(h_, i), h_, "compute_force");
    if (mirror_host) download_force(i);
}
void GradientFlow::compute_w1() {
    qcbf_compat_check(qcbf_compute_w1(h_, epsilon), h_, "compute_w1");
    if (mirror_host) download_field_c();
}
void GradientFlow::compute_w2() {
    qcbf_compat_check(qcbf_compute_w2(h_, epsilon), h_, "compute_w2");
    if (mirror_host) download_field_c();
}
void GradientFlow::replace_force_0() {
    qcbf_compat_check(qcbf_replace_force_0(h_), h_, "replace_force_0");
    if (mirror_host) download_force(0);
}
void GradientFlow::compute_w3() {
    qcbf_compat_check(qcbf_compute_w3(h_, epsilon), h_, "compute_w3");
    if (mirror_host) download_field_c();
}

// rk3_step (src/flow/gradient_flow.cpp:124-135): three fused stage kernels;
// the halo exchanges have no counterpart on one device.
void GradientFlow::rk3_step(mpi::MpiTopology&) {
    qcbf_compat_check(qcbf_rk3_step(h_, epsilon, 1), h_, "rk3_step");
    if (mirror_host) {
        download_field_c();
        field_c.fill_halos(*geo_p);
    }
}
