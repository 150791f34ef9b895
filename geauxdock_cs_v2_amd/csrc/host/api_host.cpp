// Host half of the C ABI: errors, random streams, loader and staging.
#include <cstddef>
#include <cstring>
#include <string>

#include "gd/geauxdock_amd.h"
#include "job.hpp"
#include "philox.hpp"

struct gd_raw_handle {
    gd::RawJob job;
};

struct gd_complex_handle {
    gd::StagedJob job;
};

namespace gd {
thread_local std::string g_last_error;

int fail(int code, const std::string &msg)
{
    g_last_error = msg;
    return code;
}
} // namespace gd

extern "C" {

const char *gd_last_error(void) { return gd::g_last_error.c_str(); }

const char *gd_version(void) { return "geauxdock-amd 0.1 (gfx950)"; }

void gd_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    gd::philox4x32_10(ctr, key, out);
}

uint32_t gd_stream_k(uint64_t seed, uint32_t kind, uint32_t r, uint32_t s, uint32_t j)
{
    return gd::stream_k(seed, kind, r, s, j);
}

int gd_load_job_line(const char *csv_line, gd_raw_handle **out)
{
    if (!csv_line || !out) return gd::fail(GD_ERR_ARG, "null argument");
    gd::LoadArgs args;
    std::string err;
    if (!gd::parse_job_line(csv_line, args, err)) return gd::fail(GD_ERR_ARG, err);
    gd_raw_handle *h = new gd_raw_handle();
    if (!gd::load_raw(args, h->job, err)) {
        delete h;
        return gd::fail(GD_ERR_IO, err);
    }
    *out = h;
    return GD_OK;
}

int gd_load_job(const char *prt_path, const char *lig_path, const char *ff_path, const char *lig_id,
                const char *weight_path, const char *enepara_path, const char *nora_path,
                const char *norb_path, int n_tmp, float temp_high, float temp_low,
                float tras_scale, float rot_scale, int n_dump, gd_raw_handle **out)
{
    if (!prt_path || !lig_path || !ff_path || !lig_id || !weight_path || !enepara_path || !nora_path ||
        !norb_path || !out)
        return gd::fail(GD_ERR_ARG, "null argument");
    gd::LoadArgs a;
    a.prt_path = prt_path;
    a.lig_path = lig_path;
    a.ff_path = ff_path;
    a.lig_id = lig_id;
    a.weight_path = weight_path;
    a.enepara_path = enepara_path;
    a.nora_path = nora_path;
    a.norb_path = norb_path;
    a.n_tmp = n_tmp;
    a.temp_high = temp_high;
    a.temp_low = temp_low;
    a.tras_scale = tras_scale;
    a.rot_scale = rot_scale;
    a.n_dump = n_dump;
    gd_raw_handle *h = new gd_raw_handle();
    std::string err;
    if (!gd::load_raw(a, h->job, err)) {
        delete h;
        return gd::fail(GD_ERR_IO, err);
    }
    *out = h;
    return GD_OK;
}

const gd_Raw *gd_raw_view(const gd_raw_handle *h) { return h ? &h->job.view : nullptr; }

void gd_raw_free(gd_raw_handle *h) { delete h; }

int gd_stage(const gd_Raw *raw, gd_complex_handle **out)
{
    if (!raw || !out) return gd::fail(GD_ERR_ARG, "null argument");
    gd_complex_handle *h = new gd_complex_handle();
    std::string err;
    if (!gd::stage(*raw, h->job, err)) {
        delete h;
        return gd::fail(GD_ERR_ARG, err);
    }
    *out = h;
    return GD_OK;
}

const gd_Complex *gd_complex_view(const gd_complex_handle *h) { return h ? &h->job.view : nullptr; }

void gd_complex_free(gd_complex_handle *h) { delete h; }

size_t gd_abi_layout(int which)
{
    switch (which) {
    case 0: return sizeof(gd_Complex);
    case 1: return offsetof(gd_Complex, enepara);
    case 2: return sizeof(gd_Raw);
    case 3: return offsetof(gd_Raw, enepara0);
    case 4: return sizeof(gd_EnePara);
    case 5: return sizeof(gd_EnePara0);
    case 6: return sizeof(gd_ReplicaMC);
    case 7: return sizeof(gd_Temp);
    case 8: return offsetof(gd_Complex, mcpara);
    case 9: return offsetof(gd_Raw, n_tmp);
    default: return 0;
    }
}

double gd_superposed_rmsd(const float *ax, const float *ay, const float *az, const float *bx,
                          const float *by, const float *bz, int n)
{
    if (n <= 0) return 0.0;
    return gd::superposed_rmsd(ax, ay, az, bx, by, bz, n);
}

} // extern "C"
