/*
 * ORACLE — test infrastructure only (see bf_oracle.h).
 * Exported entry points of the closure library (bf_oracle_lobes.h, bf_oracle_microfacet.h) for
 * the closed-form and invariant checks in tests/. Transport mode is radiance throughout.
 */
#include "bf_oracle.h"
#include "bf_oracle_lobes.h"

static v3 vec(const float* d) { return v3_make(d[0], d[1], d[2]); }

int bfo_closure_eval(const bf_closure_record* c, uint32_t allowed, const float wo[3], const float wi[3], float out[4]) {
    closure_rec r;
    memcpy(&r, c, sizeof r);
    return lobe_record_eval(&r, allowed, vec(wo), vec(wi), out);
}

int bfo_closure_pdf(const bf_closure_record* c, uint32_t allowed, const float wo[3], const float wi[3], float* pdf) {
    closure_rec r;
    memcpy(&r, c, sizeof r);
    return lobe_record_pdf(&r, allowed, vec(wo), vec(wi), pdf);
}

/* out: value[4], incoming[3], pdf, eta_scale; returns the status, *lobes the sampled lobe bits */
int bfo_closure_sample(const bf_closure_record* c, uint32_t allowed, const float wo[3], float event, float u, float v,
                       float out[9], uint32_t* lobes) {
    closure_rec r;
    memcpy(&r, c, sizeof r);
    lobe_sample s;
    lobe_record_sample(&r, allowed, vec(wo), event, u, v, &s);
    memcpy(out, s.value, 4 * sizeof(float));
    out[4] = s.incoming.x;
    out[5] = s.incoming.y;
    out[6] = s.incoming.z;
    out[7] = s.pdf;
    out[8] = s.eta_scale;
    *lobes = s.lobes;
    return (int)s.status;
}

int bfo_dielectric_fresnel(float cos_i, float eta_i, float eta_t, float* out) {
    return mf_dielectric_fresnel(cos_i, eta_i, eta_t, out);
}
int bfo_conductor_fresnel(float cos_i, float eta, float k, float* out) { return mf_conductor_fresnel_lane(cos_i, eta, k, out); }
int bfo_ggx_d(float ax, float ay, const float m[3], float* out) { return mf_ggx_d(ax, ay, vec(m), out); }
int bfo_smith_g1(float ax, float ay, const float d[3], float* out) { return mf_smith_g1(ax, ay, vec(d), out); }
int bfo_smith_g2(float ax, float ay, const float wo[3], const float wi[3], float* out) {
    return mf_smith_g2(ax, ay, vec(wo), vec(wi), out);
}
int bfo_smith_lambda(float ax, float ay, const float d[3], float* out) { return mf_smith_lambda(ax, ay, vec(d), out); }
int bfo_vndf_pdf(float ax, float ay, const float wo[3], const float m[3], float* out) {
    return mf_vndf_pdf(ax, ay, vec(wo), vec(m), out);
}
int bfo_sample_vndf(float ax, float ay, const float wo[3], float u, float v, float m[3], float* pdf) {
    v3 r;
    const int st = mf_sample_vndf(ax, ay, vec(wo), u, v, &r, pdf);
    m[0] = r.x;
    m[1] = r.y;
    m[2] = r.z;
    return st;
}
int bfo_ggx_widths_representable(float ax, float ay) { return mf_ggx_widths(ax, ay); }
float bfo_hypotf(float x, float y) { return bf_hypotf(x, y); }
float bfo_atan2f(float y, float x) { return bf_atan2f(y, x); }
void bfo_sincosf(float a, float* s, float* c) { bf_sincosf(a, s, c); }
float bfo_expf(float x) { return bf_expf(x); }
float bfo_expm1f(float x) { return bf_expm1f(x); }
float bfo_logf(float x) { return bf_logf(x); }

int bfo_mis_weight(uint32_t heuristic, float sampled, float competing, float* weight) {
    return lobe_mis_weight(heuristic, sampled, competing, weight);
}

int bfo_russian_roulette(float throughput[4], float eta_scale, float u, float minimum, float maximum, int* survived) {
    float p;
    return lobe_russian_roulette(throughput, eta_scale, u, minimum, maximum, survived, &p);
}
