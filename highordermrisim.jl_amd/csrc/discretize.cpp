// discretize.cpp -- block sequence description -> node grid (host code).
//
// Follows the structure of the reference's discretize
// (src/datatypes/HO_DiscreteSequence.jl:93-179): per block, the sampling times
// are the block edges, every waveform knot, every ADC time and the RF window
// edges, then filled evenly so no interval is longer than dt_max (dt_rf while
// RF is on); every waveform (gradients, B1, the sixteen field coefficients) is
// sampled on that grid; blocks are appended with the boundary rule of
// append_high_order_block! (HO_DiscreteSequence.jl:95-132): a boundary sample
// is shared when every value agrees, otherwise both sides are kept at the same
// time with a zero-length, non-RF interval between them.  The same rule keeps
// jumps inside a block (repeated knot times).  The reference delegates the
// sampling rules to KomaMRIBase (not in the reference tree), so node placement
// beyond these rules is this library's own.
#include "hom_internal.h"

#include <algorithm>
#include <cmath>
#include <new>
#include <vector>

namespace {

constexpr int kWaves = 3 + 2 + 16;  // gx gy gz, B1 re/im, h0..h15

// Piecewise-linear waveform sampled at nondecreasing times: value just
// before (left) or just after (right) time t; zero outside its knots.
// Repeated knot times are jumps.  The cursor only moves forward, so sampling
// a block costs O(knots + samples).
struct Cursor {
    const hom_wave *w;
    int64_t k = 0;  // first knot with time >= t (t = last queried time)

    double value(double t, bool right) {
        const int64_t n = w->n;
        if (n <= 0 || !w->t || !w->a) return 0.0;
        const double *kt = w->t, *ka = w->a;
        while (k < n && kt[k] < t) ++k;
        if (t < kt[0] || t > kt[n - 1]) return 0.0;
        if (right) {
            int64_t j = k;  // first knot with time > t
            while (j < n && kt[j] <= t) ++j;
            if (j >= n) return 0.0;          // at or after the last knot
            if (j == 0) return 0.0;
            const double t0 = kt[j - 1], t1 = kt[j];
            if (t == t0) return ka[j - 1];
            return ka[j - 1] + (t - t0) / (t1 - t0) * (ka[j] - ka[j - 1]);
        }
        if (k == 0) return 0.0;              // at the first knot: nothing before it
        const double t0 = kt[k - 1], t1 = kt[k];
        if (t == t1) return ka[k];           // left limit at a knot: the first value there
        return ka[k - 1] + (t - t0) / (t1 - t0) * (ka[k] - ka[k - 1]);
    }
};

struct Node {
    double t;
    double v[kWaves];
    double df, psi;
    bool adc;
};

struct Builder {
    std::vector<Node> nodes;
    std::vector<uint8_t> exc;  // per interval
};

bool same_values(const Node &a, const Node &b) {
    for (int i = 0; i < kWaves; ++i)
        if (a.v[i] != b.v[i]) return false;
    return a.df == b.df && a.psi == b.psi;
}

int check_wave(const hom_wave &w) {
    if (w.n < 0 || (w.n > 0 && (!w.t || !w.a))) return HOM_ERR_ARGUMENT;
    for (int64_t i = 0; i < w.n; ++i)
        if (!std::isfinite(w.t[i]) || !std::isfinite(w.a[i]) || (i > 0 && w.t[i] < w.t[i - 1]))
            return HOM_ERR_ARGUMENT;
    return HOM_OK;
}

int append_block(Builder &B, const hom_block &blk, double t0, double dt_max, double dt_rf,
                 bool nominal_in_h) {
    const double dur = blk.duration;
    const hom_wave *w[kWaves] = {&blk.gx, &blk.gy, &blk.gz, &blk.rf_re, &blk.rf_im};
    for (int j = 0; j < 16; ++j) w[5 + j] = &blk.h[j];
    const bool has_rf = blk.rf_t1 > blk.rf_t0;
    // key times
    std::vector<double> key = {0.0, dur};
    for (int i = 0; i < kWaves; ++i) {
        bool seen = false;  // waveforms often share one knot array
        for (int j = 0; j < i && !seen; ++j) seen = w[j]->t == w[i]->t && w[j]->n == w[i]->n;
        if (seen) continue;
        for (int64_t k = 0; k < w[i]->n; ++k)
            if (w[i]->t[k] > 0.0 && w[i]->t[k] < dur) key.push_back(w[i]->t[k]);
    }
    for (int64_t k = 0; k < blk.adc_n; ++k) key.push_back(blk.adc_t[k]);
    if (has_rf) {
        key.push_back(std::max(0.0, blk.rf_t0));
        key.push_back(std::min(dur, blk.rf_t1));
    }
    std::sort(key.begin(), key.end());
    key.erase(std::unique(key.begin(), key.end()), key.end());
    // maximum step fill; only key times can hold a jump
    std::vector<double> times;
    std::vector<uint8_t> is_key;
    times.reserve(key.size() * 2);
    is_key.reserve(key.size() * 2);
    for (size_t i = 0; i + 1 < key.size(); ++i) {
        const double lo = key[i], hi = key[i + 1];
        times.push_back(lo);
        is_key.push_back(1);
        const bool rf_on = has_rf && lo >= blk.rf_t0 && hi <= blk.rf_t1;
        const double step = rf_on ? dt_rf : dt_max;
        if (step > 0 && hi - lo > step) {
            const int64_t k = (int64_t)std::ceil((hi - lo) / step - 1e-9);
            for (int64_t j = 1; j < k; ++j) {
                times.push_back(lo + (hi - lo) * (double)j / (double)k);
                is_key.push_back(0);
            }
        }
    }
    times.push_back(key.back());
    is_key.push_back(1);
    std::vector<double> adc(blk.adc_t, blk.adc_t + blk.adc_n);
    std::sort(adc.begin(), adc.end());
    Cursor cur[kWaves];
    for (int i = 0; i < kWaves; ++i) cur[i].w = w[i];
    auto sample = [&](double t, bool right) {
        Node n;
        n.t = t0 + t;
        const bool in_rf = has_rf && (right ? (t >= blk.rf_t0 && t < blk.rf_t1) : (t > blk.rf_t0 && t <= blk.rf_t1));
        for (int i = 0; i < kWaves; ++i) n.v[i] = cur[i].value(t, right);
        if (!in_rf) n.v[3] = n.v[4] = 0.0;
        if (nominal_in_h)
            for (int j = 0; j < 3; ++j) n.v[5 + 1 + j] += n.v[j];
        n.df = in_rf ? blk.rf_df : 0.0;
        n.psi = in_rf ? blk.rf_phase : 0.0;
        n.adc = std::binary_search(adc.begin(), adc.end(), t);
        return n;
    };
    for (size_t i = 0; i < times.size(); ++i) {
        const double t = times[i];
        const bool first = i == 0, last = i + 1 == times.size();
        Node here[2];
        int nh = 1;
        if (first) here[0] = sample(t, true);
        else if (last || !is_key[i]) here[0] = sample(t, false);  // continuous away from knots
        else {
            here[0] = sample(t, false);
            here[1] = sample(t, true);
            if (same_values(here[0], here[1])) here[0] = here[1];
            else {
                here[1].adc = false;  // the ADC sample sits on the first node of a jump
                nh = 2;
            }
        }
        for (int h = 0; h < nh; ++h) {
            Node &n = here[h];
            if (B.nodes.empty()) {
                B.nodes.push_back(n);
                continue;
            }
            const Node &prev = B.nodes.back();
            if (first && h == 0 && n.t == prev.t && same_values(prev, n)) {
                B.nodes.back().adc = prev.adc || n.adc;  // shared block boundary
                continue;
            }
            const double a = prev.t, b = n.t;
            const bool rf_interval = has_rf && b > a && (a - t0) >= blk.rf_t0 && (b - t0) <= blk.rf_t1 &&
                                     !(first && h == 0);
            B.exc.push_back(rf_interval ? 1 : 0);
            B.nodes.push_back(n);
        }
    }
    return HOM_OK;
}

}  // namespace

struct hom_discrete {
    std::vector<Node> nodes;
    std::vector<uint8_t> exc;
};

extern "C" {

static int hom_discretize_impl(const hom_block *blocks, int64_t n_blocks, double dt_max, double dt_rf,
                               int nominal_in_h, hom_discrete **out, int64_t *n_nodes) {
    if (!out || !n_nodes || (n_blocks > 0 && !blocks) || n_blocks < 0) return HOM_ERR_ARGUMENT;
    *out = nullptr;
    if (!(dt_max > 0.0) || !(dt_rf > 0.0)) return HOM_ERR_ARGUMENT;
    for (int64_t b = 0; b < n_blocks; ++b) {
        const hom_block &k = blocks[b];
        if (!(k.duration >= 0.0) || !std::isfinite(k.duration)) return HOM_ERR_ARGUMENT;
        const hom_wave *w[kWaves] = {&k.gx, &k.gy, &k.gz, &k.rf_re, &k.rf_im};
        for (int j = 0; j < 16; ++j) w[5 + j] = &k.h[j];
        for (int i = 0; i < kWaves; ++i)
            if (check_wave(*w[i])) return HOM_ERR_ARGUMENT;
        if (k.adc_n < 0 || (k.adc_n > 0 && !k.adc_t)) return HOM_ERR_ARGUMENT;
        for (int64_t i = 0; i < k.adc_n; ++i)
            if (!(k.adc_t[i] > 0.0 && k.adc_t[i] <= k.duration)) return HOM_ERR_ARGUMENT;
    }
    Builder B;
    double t0 = 0.0;
    for (int64_t b = 0; b < n_blocks; ++b) {
        const int rc = append_block(B, blocks[b], t0, dt_max, dt_rf, nominal_in_h != 0);
        if (rc) return rc;
        t0 += blocks[b].duration;
    }
    hom_discrete *d = new (std::nothrow) hom_discrete();
    if (!d) return HOM_ERR_MEMORY;
    d->nodes.swap(B.nodes);
    d->exc.swap(B.exc);
    *n_nodes = (int64_t)d->nodes.size();
    *out = d;
    return HOM_OK;
}

static int hom_discrete_copy_impl(const hom_discrete *d, double *t, double *gx, double *gy, double *gz,
                                  double *b1, double *df, double *psi, uint8_t *adc, uint8_t *excitation,
                                  double *h) {
    if (!d) return HOM_ERR_ARGUMENT;
    const int64_t n = (int64_t)d->nodes.size();
    for (int64_t i = 0; i < n; ++i) {
        const Node &nd = d->nodes[(size_t)i];
        if (t) t[i] = nd.t;
        if (gx) gx[i] = nd.v[0];
        if (gy) gy[i] = nd.v[1];
        if (gz) gz[i] = nd.v[2];
        if (b1) {
            b1[2 * i] = nd.v[3];
            b1[2 * i + 1] = nd.v[4];
        }
        if (df) df[i] = nd.df;
        if (psi) psi[i] = nd.psi;
        if (adc) adc[i] = nd.adc ? 1 : 0;
        if (h)
            for (int j = 0; j < 16; ++j) h[(size_t)j * (size_t)n + (size_t)i] = nd.v[5 + j];
    }
    if (excitation)
        for (int64_t i = 0; i + 1 < n; ++i) excitation[i] = d->exc[(size_t)i];
    return HOM_OK;
}

int hom_discrete_free(hom_discrete *d) {
    delete d;
    return HOM_OK;
}

// Entry points that allocate host memory or start threads: no exception
// crosses the C ABI.
int hom_discretize(const hom_block *blocks, int64_t n_blocks, double dt_max, double dt_rf,
                   int nominal_in_h, hom_discrete **out, int64_t *n_nodes) {
    return hom::guarded_host([&] {
        return hom_discretize_impl(blocks, n_blocks, dt_max, dt_rf, nominal_in_h, out, n_nodes);
    });
}

int hom_discrete_copy(const hom_discrete *d, double *t, double *gx, double *gy, double *gz,
                      double *b1, double *df, double *psi, uint8_t *adc, uint8_t *excitation,
                      double *h) {
    return hom::guarded_host([&] {
        return hom_discrete_copy_impl(d, t, gx, gy, gz, b1, df, psi, adc, excitation, h);
    });
}

}  // extern "C"
