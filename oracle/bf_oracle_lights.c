/*
 * ORACLE (test infrastructure only): the power-weighted light sampler of the reference renderer,
 * restated in C. Renderer/Source/LightSampler.cpp: power_weights (:201-249), build_power_cdf
 * (:116-199) with add_compensated / compensated_sum (:85-114), create_power_weighted (:605-624);
 * sample (:1023-1057) and probability (:1133-1151) are used by the transport. Registry order:
 * punctual lights, then mesh-area lights (the scene's registration order).
 *
 * Statuses: 0 ok, BF_INVALID_ARGUMENT (the reference's invalid_argument), BF_RESOURCE_EXHAUSTED
 * (resource_exhausted), BF_INTERNAL_ERROR (internal_error).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "bf_oracle.h"
#include "bf_oracle_internal.h"

/* Neumaier step on (sum, c); 0 when the term is negative / not finite or a partial overflows. */
static int neumaier(float* sum, float* c, float x) {
    if (!(x >= 0.0f) || isinf(x)) return 0;
    const float t = *sum + x;
    if (isinf(t)) return 0;
    if (*sum >= x) *c += (*sum - t) + x;
    else *c += (x - t) + *sum;
    if (!isfinite(*c)) return 0;
    *sum = t;
    return 1;
}

int bfo_light_power_cdf(const float* powers, uint32_t n, float* cdf) {
    if (n == 0 || !powers || !cdf) return BF_INVALID_ARGUMENT;
    float peak = 0.0f;
    for (uint64_t k = 0; k < 4ull * n; ++k) {
        const float v = powers[k];
        if (!isfinite(v) || v < 0.0f) return BF_INVALID_ARGUMENT;
        if (v > peak) peak = v;
    }
    if (peak == 0.0f) return BF_INVALID_ARGUMENT;
    float* w = calloc(n, sizeof *w);
    if (!w) return BF_RESOURCE_EXHAUSTED;
    int status = 0;
    for (uint32_t i = 0; i < n && !status; ++i) {
        float s = 0.0f, c = 0.0f;
        int any = 0;
        for (int k = 0; k < 4; ++k) {
            const float lane = powers[4ull * i + k];
            const float r = lane / peak;
            if (lane > 0.0f) {
                any = 1;
                if (!(r > 0.0f)) status = BF_RESOURCE_EXHAUSTED;
            }
            if (!status && !neumaier(&s, &c, r)) status = BF_RESOURCE_EXHAUSTED;
        }
        if (status) break;
        const float v = s + c;
        if (!isfinite(v) || v < 0.0f || (any && !(v > 0.0f))) status = BF_RESOURCE_EXHAUSTED;
        else w[i] = v == 0.0f ? 0.0f : v; /* -0 -> +0 */
    }
    float total_s = 0.0f, total_c = 0.0f, total = 0.0f;
    if (!status) {
        for (uint32_t i = 0; i < n && !status; ++i)
            if (!neumaier(&total_s, &total_c, w[i])) status = BF_RESOURCE_EXHAUSTED;
        total = total_s + total_c;
        if (!status && (!isfinite(total) || total < 0.0f)) status = BF_RESOURCE_EXHAUSTED;
        if (!status && total == 0.0f) status = BF_INVALID_ARGUMENT;
    }
    int64_t last = -1;
    for (int64_t i = (int64_t)n - 1; i >= 0 && !status; --i)
        if (w[i] > 0.0f) {
            last = i;
            break;
        }
    if (!status && last < 0) status = BF_INTERNAL_ERROR;
    if (!status) {
        float ps = 0.0f, pc = 0.0f, prev = 0.0f;
        cdf[0] = 0.0f;
        for (uint32_t i = 0; i < n && !status; ++i) {
            if (w[i] == 0.0f) {
                cdf[i + 1] = prev;
                continue;
            }
            if (!neumaier(&ps, &pc, w[i])) {
                status = BF_RESOURCE_EXHAUSTED;
                break;
            }
            const float prefix = ps + pc;
            if (!isfinite(prefix) || !(prefix > 0.0f)) {
                status = BF_RESOURCE_EXHAUSTED;
                break;
            }
            float b = 1.0f;
            if ((int64_t)i == last) {
                const float nominal = w[i] / total;
                const float after = prev + nominal;
                if (!isfinite(nominal) || !(nominal > 0.0f) || !isfinite(after) || !(after > prev)) status = BF_RESOURCE_EXHAUSTED;
            } else {
                b = prefix / total;
                if (!isfinite(b) || !(b > prev) || !(b < 1.0f)) status = BF_RESOURCE_EXHAUSTED;
            }
            cdf[i + 1] = b;
            prev = b;
        }
    }
    if (!status && cdf[n] != 1.0f) status = BF_INTERNAL_ERROR;
    for (uint32_t i = 0; i < n && !status; ++i) {
        const float p = cdf[i + 1] - cdf[i];
        if (!isfinite(p) || p < 0.0f || ((w[i] > 0.0f) != (p > 0.0f))) status = BF_RESOURCE_EXHAUSTED;
    }
    free(w);
    return status;
}

int bfo_scene_set_light_powers(bfo_scene* scene, const float* powers, uint32_t count) {
    if (!scene) return BF_INVALID_ARGUMENT;
    if (!powers) {
        free(scene->light_cdf);
        scene->light_cdf = NULL;
        return 0;
    }
    if (count != scene->punctual_count + scene->light_count) return BF_INVALID_ARGUMENT;
    float* cdf = malloc(((size_t)count + 1) * sizeof *cdf);
    if (!cdf) return BF_RESOURCE_EXHAUSTED;
    const int status = bfo_light_power_cdf(powers, count, cdf);
    if (status) {
        free(cdf);
        return status;
    }
    free(scene->light_cdf);
    scene->light_cdf = cdf;
    return 0;
}

/* sample (:1023-1057): the slot whose interval [cdf[i], cdf[i + 1]) holds u, by upper_bound */
int bfo_light_cdf_sample(const float* cdf, uint32_t n, float u, uint32_t* index, float* probability) {
    if (n == 0 || !isfinite(u) || u < 0.0f || !(u < 1.0f)) return 0;
    const float* it = cdf + 1;
    size_t len = n;
    while (len > 0) { /* first element of cdf[1..n] greater than u */
        const size_t half = len / 2;
        if (it[half] <= u) {
            it += half + 1;
            len -= half + 1;
        } else {
            len = half;
        }
    }
    const size_t offset = (size_t)(it - cdf);
    if (offset > n) return 0;
    const float p = cdf[offset] - cdf[offset - 1];
    if (!isfinite(p) || !(p > 0.0f) || p > 1.0f) return 0;
    *index = (uint32_t)(offset - 1);
    *probability = p;
    return 1;
}

/* probability (:1133-1151) */
int bfo_light_cdf_probability(const float* cdf, uint32_t n, uint32_t i, float* probability) {
    if (i >= n) return 0;
    const float p = cdf[i + 1] - cdf[i];
    if (!isfinite(p) || p < 0.0f || p > 1.0f) return 0;
    *probability = p;
    return 1;
}
