// gradient_flow.h -- drop-in for src/flow/gradient_flow.h:11-27.
//
// Same members and methods as the reference's GradientFlow; the flowing field
// lives in HBM (a libqcbflow handle) and every method runs HIP kernels.  The
// public GaugeField members are host mirrors: with mirror_host = true (the
// default) each method that changes the device state refreshes them, so code
// that reads gf.field_c after rk3_step sees what the reference would show.
// Set mirror_host = false to skip those PCIe copies when only the observables
// are needed (topo_charge_flowed refreshes field_c once at its end either way).
// Errors from the device abort the process with status 1, as the reference's
// MPI_Abort paths do.
#pragma once
#include <vector>

#include "../gauge/GaugeField.h"
#include "../mpi/MpiTopology.h"
#include "qcbflow.h"

class GradientFlow {
public:
    double epsilon;
    GaugeField field_c;
    GaugeField force_0;
    GaugeField force_1;
    const GeometryCB* geo_p;
    bool mirror_host = true;

    explicit GradientFlow(double epsilon_, const GaugeField& field, const GeometryCB& geo);
    ~GradientFlow();
    GradientFlow(const GradientFlow&) = delete;
    GradientFlow& operator=(const GradientFlow&) = delete;

    void copy(const GaugeField& field);
    void compute_force(int i);
    void compute_w1();
    void compute_w2();
    void replace_force_0();
    void compute_w3();
    void rk3_step(mpi::MpiTopology& topo);

    // The device handle (for callers that use the C ABI directly) and an
    // explicit refresh of field_c from the device.
    [[nodiscard]] qcbf_flow* device() const { return h_; }
    void download_field_c();

private:
    qcbf_flow* h_ = nullptr;
    std::vector<double> buf_;
    void download_force(int i);
};

// Abort with the library's message when a C-ABI call fails (used by the shim).
void qcbf_compat_check(int rc, const qcbf_flow* h, const char* what);
