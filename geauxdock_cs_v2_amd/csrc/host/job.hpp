// Host-side owners of the loader output (RawJob, gd_Raw view) and of the
// staged tables (StagedJob, gd_Complex view).
#pragma once

#include <string>
#include <vector>

#include "gd/gd_types.h"

namespace gd {

// Input of one docking job: the columns of a reference CSV job line
// (src/common/size.h:136-161, src/frontend/frontend.cpp:70-123).
struct LoadArgs {
    int data_id = 0;
    std::string prt_id, prt_path;
    std::string lig_id, lig_path, ff_path;
    std::string weight_path, enepara_path, nora_path, norb_path;
    int n_tmp = 1;
    float temp_high = 0.044f, temp_low = 0.036f;
    float tras_scale = 0.02f, rot_scale = 0.08f;
    int n_dump = 1;
    int max_conf_lig = 64; // MAX_CONF_LIG (src/common/size.h:37)
};

struct RawJob {
    gd_Raw view{};
    std::vector<float> lig_x, lig_y, lig_z, lig_c;
    std::vector<int> lig_t;
    std::vector<float> prt_x, prt_y, prt_z;
    std::vector<int> prt_t, prt_c, prt_d, prt_r, prt_seq3;
    std::vector<int> psp_res, psp_lig;
    std::vector<float> psp_val;
    std::vector<int> kde_t;
    std::vector<float> kde_x, kde_y, kde_z;
    std::vector<float> mcs_tcc, mcs_x, mcs_y, mcs_z;
    std::vector<int> mcs_rowptr, mcs_i;
    std::string lig_name;
    void bind(); // point view at the vectors
};

struct StagedJob {
    gd_Complex view{};
    std::vector<float> lig_x2, lig_y2, lig_z2, lig_c2, lig_x3, lig_y3, lig_z3, lig_c3;
    std::vector<int> lig_t2, lig_t3, lig_kde_begin_idx, lig_kde_end_idx;
    std::vector<float> prt_x, prt_y, prt_z, prt_hpp, prt_ele;
    std::vector<int> prt_t, prt_c, prt_cdc, prt_seq3r;
    std::vector<float> psp;
    std::vector<int> kde_t;
    std::vector<float> kde_x, kde_y, kde_z;
    std::vector<int> mcs_ncol, mcs_rowptr, mcs_i;
    std::vector<float> mcs_tcc, mcs_x, mcs_y, mcs_z;
    std::vector<gd_Temp> temp;
    std::vector<gd_ReplicaMC> replica;
    void bind();
};

bool parse_job_line(const std::string &line, LoadArgs &out, std::string &err);
bool load_raw(const LoadArgs &args, RawJob &out, std::string &err);
bool stage(const gd_Raw &raw, StagedJob &out, std::string &err);

// Pieces of the loader, exposed for tests.
bool load_ligand(const std::string &path, int max_conf, RawJob &out, std::string &err);
bool load_protein(const std::string &path, RawJob &out, std::string &err);
bool load_lhm(const std::string &path, const std::string &lig_id, RawJob &out, std::string &err);
bool load_enepara(const std::string &path, gd_EnePara0 &ep, std::string &err);
bool load_column_file(const std::string &path, float *dst, int cap, std::string &err);

// Best-superposition RMSD between two centred coordinate sets (Kabsch /
// Horn quaternion), the role of u3b in src/frontend/rmsd.f:16-255.
double superposed_rmsd(const float *ax, const float *ay, const float *az,
                       const float *bx, const float *by, const float *bz, int n);

} // namespace gd
