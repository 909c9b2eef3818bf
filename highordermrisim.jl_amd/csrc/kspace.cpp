// kspace.cpp -- encoding coefficients of a discretized sequence (host code).
//
// Restates grad2kspace (reference src/datatypes/HO_Sequence.jl:256-287) on the
// node grid of an HO_DiscreteSequence: per field term, the trapezoid integral
// of the waveform from the start of its part, where parts are separated by RF
// events; entering a part, the integral continues from the previous part's
// end times +1 (skip_rf or neither kind), 0 (excitation) or -1 (refocusing);
// k = gamma * integral.  Row i belongs to the end of interval i (time
// t[i+1]).  ADC rows take the value at their node.
//
// The reference reads the RF kinds and break points from KomaMRIBase (not in
// the reference tree).  Here an RF event is a maximal run of excitation
// intervals; its flip angle is 360 * gamma * sum |B1[n]| dt[n] degrees (the
// left-point rule of the excitation kernel); flip <= 90.01 degrees is an
// excitation, larger a refocusing pulse; the part break is the first interval
// that starts at or after the event's mid time.
#include "hom_internal.h"

#include <cmath>
#include <cstring>
#include <vector>

namespace {

struct RfEvent {
    int64_t first, last;  // intervals [first, last)
    double flip_deg;
    int64_t brk;          // first interval of the new part
};

std::vector<RfEvent> rf_events(const hom_sequence_host *s, double gamma) {
    std::vector<RfEvent> ev;
    const int64_t ni = s->n_nodes - 1;
    for (int64_t a = 0; a < ni;) {
        if (!s->excitation[a]) {
            ++a;
            continue;
        }
        int64_t b = a;
        double area = 0.0;
        while (b < ni && s->excitation[b]) {
            area += std::hypot(s->b1[2 * b], s->b1[2 * b + 1]) * s->dt[b];
            ++b;
        }
        RfEvent e{a, b, 360.0 * gamma * area, b};
        const double mid = 0.5 * (s->t[a] + s->t[b]);
        for (int64_t i = a; i < b; ++i)
            if (s->t[i] >= mid) {
                e.brk = i;
                break;
            }
        ev.push_back(e);
        a = b;
    }
    return ev;
}

}  // namespace

extern "C" {

static int hom_kspace_rf_events_impl(const hom_sequence_host *seq, double gamma, int64_t *n_events,
                                     double *flip_deg, int64_t capacity) {
    if (!seq || !n_events) return HOM_ERR_ARGUMENT;
    if (seq->n_nodes < 1) return HOM_ERR_DIMENSION;
    if (seq->n_nodes > 1 && (!seq->t || !seq->dt || !seq->b1 || !seq->excitation))
        return HOM_ERR_ARGUMENT;
    const std::vector<RfEvent> ev = rf_events(seq, gamma);
    *n_events = (int64_t)ev.size();
    if (flip_deg)
        for (int64_t i = 0; i < (int64_t)ev.size() && i < capacity; ++i) flip_deg[i] = ev[(size_t)i].flip_deg;
    return HOM_OK;
}

static int hom_kspace_impl(const hom_sequence_host *seq, double gamma, const uint8_t *skip_rf, int64_t n_skip,
                           double *k_nominal, double *k_nominal_adc, double *k_dfc, double *k_dfc_adc) {
    if (!seq) return HOM_ERR_ARGUMENT;
    const int64_t n = seq->n_nodes, ni = n - 1;
    if (ni < 1) return HOM_ERR_ARGUMENT;  // "Cannot integrate an empty sequence."
    if (!seq->t || !seq->dt || !seq->gx || !seq->gy || !seq->gz || !seq->b1 || !seq->adc ||
        !seq->excitation)
        return HOM_ERR_ARGUMENT;
    for (int j = 0; j < 16; ++j)
        if (!seq->h[j]) return HOM_ERR_ARGUMENT;
    const std::vector<RfEvent> ev = rf_events(seq, gamma);
    if (skip_rf && n_skip != (int64_t)ev.size()) return HOM_ERR_DIMENSION;
    // part start flags and the sign applied to the previous part's end
    std::vector<int> sign_at((size_t)ni, 2);  // 2: no break here
    for (size_t e = 0; e < ev.size(); ++e) {
        const bool skip = skip_rf && skip_rf[e];
        sign_at[(size_t)ev[e].brk] = skip ? 1 : (ev[e].flip_deg <= 90.01 ? 0 : -1);
    }
    const double *w[19] = {seq->gx, seq->gy, seq->gz};
    for (int j = 0; j < 16; ++j) w[3 + j] = seq->h[j];
    std::vector<int64_t> adc_nodes;
    for (int64_t i = 0; i < n; ++i)
        if (seq->adc[i]) adc_nodes.push_back(i);
    const int64_t nadc = (int64_t)adc_nodes.size();
    for (int term = 0; term < 19; ++term) {
        double *out = term < 3 ? (k_nominal ? k_nominal + term * ni : nullptr)
                               : (k_dfc ? k_dfc + (term - 3) * ni : nullptr);
        double *out_adc = term < 3 ? (k_nominal_adc ? k_nominal_adc + term * nadc : nullptr)
                                   : (k_dfc_adc ? k_dfc_adc + (term - 3) * nadc : nullptr);
        std::vector<double> k((size_t)ni);
        double acc = 0.0, prev_end = 0.0;
        const double *g = w[term];
        for (int64_t i = 0; i < ni; ++i) {
            if (sign_at[(size_t)i] != 2 && i > 0) {
                prev_end = acc;
                acc = sign_at[(size_t)i] * prev_end;
            }
            acc += (g[i] + g[i + 1]) * (seq->dt[i] / 2.0);
            k[(size_t)i] = gamma * acc;
        }
        if (out) std::memcpy(out, k.data(), sizeof(double) * (size_t)ni);
        if (out_adc)
            for (int64_t m = 0; m < nadc; ++m) {
                const int64_t node = adc_nodes[(size_t)m];
                out_adc[m] = node == 0 ? 0.0 : k[(size_t)(node - 1)];
            }
    }
    return HOM_OK;
}

// Entry points that allocate host memory or start threads: no exception
// crosses the C ABI.
int hom_kspace(const hom_sequence_host *seq, double gamma, const uint8_t *skip_rf, int64_t n_skip,
               double *k_nominal, double *k_nominal_adc, double *k_dfc, double *k_dfc_adc) {
    return hom::guarded_host([&] {
        return hom_kspace_impl(seq, gamma, skip_rf, n_skip, k_nominal, k_nominal_adc, k_dfc, k_dfc_adc);
    });
}

int hom_kspace_rf_events(const hom_sequence_host *seq, double gamma, int64_t *n_events,
                         double *flip_deg, int64_t capacity) {
    return hom::guarded_host([&] {
        return hom_kspace_rf_events_impl(seq, gamma, n_events, flip_deg, capacity);
    });
}

}  // extern "C"
