// gd_dock -- the reference's docking driver over the MI355X library.
//
// Mirrors src/main/3_geauxdock.cpp: read a CSV job list (first line a
// header), and for every job line load the files it names, stage the
// complex, and run the Docker (here gd::Docker over the C ABI,
// include/gd/docker.hpp).  Per job it prints the first record entry of
// replica 0 in the PrintCsv format (src/common/util_print.cpp:77-110; for
// the 1a07C job that is the known-answer line of util_print.cpp:351) and
// writes the per-replica record dump of util_print.cpp:378-435 to
// <out_dir>/out_<lig_id>.txt.  A job that fails to load is reported and
// skipped (the reference exits instead).
//
// With --serve the job list is a request stream instead, the role of the
// reference's MPI job farm (src/main_mpi_noschedule/{server,client}.cpp: a
// server hands CSV job lines to workers): every job line becomes a request
// of the serving loop (gd_server_*), jobs against the same protein file share
// one resident target and run batched, and each job's output is its final
// replica states (the serving loop keeps no per-step records): replica 0 in
// the PrintCsv format on stdout, every replica in <out_dir>/out_<lig_id>.txt.
//
//   gd_dock <jobs.csv> [--out DIR] [--seed S] [--device D] [--serve]
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "gd/docker.hpp"

namespace {

const char *kTerms[GD_MAXWEI] = {"vdw", "ele", "pmf", "psp", "hdb", "hpc", "kde", "lhm", "dst", "total"};

void print_csv(FILE *f, int idx_rep, int step, const float *energy)
{
    std::fprintf(f, "%3d %d", idx_rep, step);
    for (int i = 0; i < GD_MAXWEI; ++i) std::fprintf(f, " %.4f", energy[i]);
    std::fprintf(f, "\n");
}

bool write_dump(const std::string &path, const char *prt_file, const char *lig_file, const gd_Complex *cx,
                const gd::Docker &d)
{
    FILE *fp = std::fopen(path.c_str(), "w");
    if (!fp) return false;
    const gd_ComplexSize &z = cx->size;
    std::fprintf(fp, "prt input file: %s, \t%3d conformations\n", prt_file, z.n_prt);
    std::fprintf(fp, "lig input file: %s, \t%3d conformations\n", lig_file, z.n_lig);
    std::fprintf(fp, "\n");
    for (int r = 0; r < z.n_rep; ++r) {
        const gd_ReplicaMC &rep = d.replica[r];
        std::fprintf(fp, "replica, %3d\n", r);
        std::fprintf(fp, "    prt  index, %3d\n", rep.idx_prt);
        std::fprintf(fp, "    temp index, %3d, minus beta, %f\n", rep.idx_tmp, cx->temp[rep.idx_tmp].minus_beta);
        std::fprintf(fp, "    lig  index, %3d \n", rep.idx_lig);
        const int ne = d.next_entry[r];
        std::fprintf(fp, "    number of accepted moves, %d\n", ne - 1);
        std::fprintf(fp, "    number of proposed moves, %d\n", cx->mcpara.steps_total - 1);
        std::fprintf(fp, "    The first record entry contains the initial status.\n");
        std::fprintf(fp, "    Records in format \"record entry, mc move step, translation x y z, rotation x y z, "
                         "evdw, eele, epmf, epsp, ehdb, ehpc, ekde, emcs, edst, etotal\"\n");
        for (int s = 0; s < ne; ++s) {
            const gd_ReplicaMC &e = d.record[(size_t)r * d.record_cap + s];
            std::fprintf(fp, "   %5d, %5d, ", s, e.step);
            for (int k = 0; k < 6; ++k) std::fprintf(fp, "%8.4f, ", e.movematrix[k]);
            for (int k = 0; k < GD_MAXWEI; ++k) std::fprintf(fp, "%8.4f, ", e.energy[k]);
            std::fprintf(fp, "\n");
        }
        std::fprintf(fp, "\n");
    }
    return std::fclose(fp) == 0;
}

bool write_final(const std::string &path, const char *prt_file, const char *lig_file, const gd_Complex *cx,
                 const std::vector<gd_ReplicaMC> &rep, const gd_request_result &res)
{
    FILE *fp = std::fopen(path.c_str(), "w");
    if (!fp) return false;
    const gd_ComplexSize &z = cx->size;
    std::fprintf(fp, "prt input file: %s, \t%3d conformations\n", prt_file, z.n_prt);
    std::fprintf(fp, "lig input file: %s, \t%3d conformations\n", lig_file, z.n_lig);
    std::fprintf(fp, "served: %d steps, latency %.1f ms (submit to result)\n", cx->mcpara.steps_total,
                 res.t_done_ms - res.t_submit_ms);
    std::fprintf(fp, "\n");
    std::fprintf(fp, "Final states in format \"replica, last accepted step, temp index, translation x y z, "
                     "rotation x y z, evdw, eele, epmf, epsp, ehdb, ehpc, ekde, emcs, edst, etotal\"\n");
    for (int r = 0; r < z.n_rep; ++r) {
        const gd_ReplicaMC &e = rep[r];
        std::fprintf(fp, "   %5d, %5d, %3d, ", r, e.step, e.idx_tmp);
        for (int k = 0; k < 6; ++k) std::fprintf(fp, "%8.4f, ", e.movematrix[k]);
        for (int k = 0; k < GD_MAXWEI; ++k) std::fprintf(fp, "%8.4f, ", e.energy[k]);
        std::fprintf(fp, "\n");
    }
    return std::fclose(fp) == 0;
}

// CSV columns of a job line (src/common/size.h:136-141): 3 protein file,
// 4 ligand id, 5 ligand file
std::vector<std::string> split_csv(const std::string &line)
{
    std::vector<std::string> out(1);
    for (char c : line) {
        if (c == ',') out.emplace_back();
        else if (c != '\r') out.back() += c;
    }
    return out;
}

struct Staged {
    size_t line = 0;
    gd_raw_handle *raw = nullptr;
    gd_complex_handle *cxh = nullptr;
    const gd_Complex *cx = nullptr;
    std::vector<std::string> col;
    int req = -1;
};

void check(int rc)
{
    if (rc != GD_OK) throw std::runtime_error(std::string("geauxdock: ") + gd_last_error());
}

// The job list as a request stream: stage every line, then per protein file
// one resident target and one serving loop for all of its jobs.
int serve_jobs(const std::vector<std::string> &lines, const std::string &out_dir, uint64_t seed, int device)
{
    int failed = 0;
    std::vector<Staged> jobs;
    std::vector<std::string> prts; // protein files in order of first use
    for (size_t i = 1; i < lines.size(); ++i) {
        Staged j;
        j.line = i;
        j.col = split_csv(lines[i]);
        if (gd_load_job_line(lines[i].c_str(), &j.raw) != GD_OK || gd_stage(gd_raw_view(j.raw), &j.cxh) != GD_OK) {
            std::fprintf(stderr, "job %zu: %s\n", i, gd_last_error());
            if (j.raw) gd_raw_free(j.raw);
            ++failed;
            continue;
        }
        j.cx = gd_complex_view(j.cxh);
        const std::string prt = j.col.size() > 3 ? j.col[3] : "";
        if (std::find(prts.begin(), prts.end(), prt) == prts.end()) prts.push_back(prt);
        jobs.push_back(std::move(j));
    }
    gd_ctx *ctx = nullptr;
    check(gd_create(device, &ctx));
    size_t hbm = 0;
    check(gd_device_info(ctx, nullptr, 0, nullptr, &hbm));
    const auto t0 = std::chrono::steady_clock::now();
    long long poses = 0;
    for (const std::string &prt : prts) {
        std::vector<Staged *> g;
        for (Staged &j : jobs)
            if ((j.col.size() > 3 ? j.col[3] : "") == prt) g.push_back(&j);
        int target = -1;
        check(gd_upload_target(ctx, g[0]->cx, &target));
        const int chunk = g[0]->cx->mcpara.steps_per_dump;
        gd_server_opts o{};
        o.hbm_budget = hbm / 4 * 3;
        o.max_active_jobs = 256;
        o.chunk_steps = chunk;
        o.exchange_interval = chunk;
        gd_server *srv = nullptr;
        check(gd_server_create(ctx, target, &o, &srv));
        for (Staged *j : g)
            if (gd_server_submit(srv, j->cx, j->cx->mcpara.steps_total, seed, &j->req) != GD_OK) {
                std::fprintf(stderr, "job %zu: %s\n", j->line, gd_last_error());
                ++failed;
            }
        check(gd_server_drain(srv));
        for (Staged *j : g) {
            if (j->req < 0) continue;
            const gd_ComplexSize &z = j->cx->size;
            std::vector<gd_ReplicaMC> rep(z.n_rep);
            gd_request_result res{};
            check(gd_server_poll(srv, j->req, &res, rep.data()));
            if (res.status != GD_OK) {
                std::fprintf(stderr, "job %zu: request failed (%d)\n", j->line, res.status);
                ++failed;
                continue;
            }
            poses += (long long)z.n_rep * j->cx->mcpara.steps_total;
            print_csv(stdout, 0, rep[0].step, rep[0].energy);
            const std::string lig_id = j->col.size() > 4 ? j->col[4] : std::to_string(j->line);
            const std::string path = out_dir + "/out_" + lig_id + ".txt";
            if (!write_final(path, prt.c_str(), j->col.size() > 5 ? j->col[5].c_str() : "", j->cx, rep, res)) {
                std::fprintf(stderr, "job %zu: cannot write %s\n", j->line, path.c_str());
                ++failed;
            }
        }
        gd_server_destroy(srv);
        check(gd_free_target(ctx, target));
    }
    const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::fprintf(stderr, "served %zu jobs: %lld pose evaluations in %.2f s (%.4g per s)\n", jobs.size(), poses, sec,
                 sec > 0 ? poses / sec : 0.0);
    gd_destroy(ctx);
    for (Staged &j : jobs) {
        gd_complex_free(j.cxh);
        gd_raw_free(j.raw);
    }
    return failed ? EXIT_FAILURE : EXIT_SUCCESS;
}

} // namespace

int main(int argc, char **argv)
{
    std::string fn, out_dir = ".";
    uint64_t seed = 2024;
    int device = 0;
    bool serve = false;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "--serve") serve = true;
        else if (a == "--out" && i + 1 < argc) out_dir = argv[++i];
        else if (a == "--seed" && i + 1 < argc) seed = std::strtoull(argv[++i], nullptr, 10);
        else if (a == "--device" && i + 1 < argc) device = std::atoi(argv[++i]);
        else if (fn.empty()) fn = a;
        else {
            std::fprintf(stderr, "%s <FILE(*.csv)> [--out DIR] [--seed S] [--device D] [--serve]\n", argv[0]);
            return EXIT_FAILURE;
        }
    }
    if (fn.empty()) fn = "data2.csv";
    std::ifstream in(fn);
    if (!in) {
        std::fprintf(stderr, "cannot open %s\n", fn.c_str());
        return EXIT_FAILURE;
    }
    std::vector<std::string> lines;
    for (std::string l; std::getline(in, l);)
        if (!l.empty()) lines.push_back(l);
    if (lines.size() < 2) {
        std::fprintf(stderr, "number of lines is wrong\n");
        return EXIT_FAILURE;
    }
    if (serve) {
        std::printf("rep step");
        for (const char *t : kTerms) std::printf(" %s", t);
        std::printf("\n");
        try {
            return serve_jobs(lines, out_dir, seed, device);
        } catch (const std::exception &e) {
            std::fprintf(stderr, "%s\n", e.what());
            return EXIT_FAILURE;
        }
    }
    int failed = 0;
    try {
        gd::Docker docker(device);
        docker.seed = seed;
        std::printf("rep step");
        for (const char *t : kTerms) std::printf(" %s", t);
        std::printf("\n");
        for (size_t i = 1; i < lines.size(); ++i) {
            gd_raw_handle *raw = nullptr;
            gd_complex_handle *cxh = nullptr;
            if (gd_load_job_line(lines[i].c_str(), &raw) != GD_OK || gd_stage(gd_raw_view(raw), &cxh) != GD_OK) {
                std::fprintf(stderr, "job %zu: %s\n", i, gd_last_error());
                if (raw) gd_raw_free(raw);
                ++failed;
                continue;
            }
            const gd_Complex *cx = gd_complex_view(cxh);
            docker.complex = cx;
            docker.run();
            const gd_ReplicaMC &first = docker.record[0];
            print_csv(stdout, 0, first.step, first.energy);
            const std::vector<std::string> col = split_csv(lines[i]);
            const std::string lig_id = col.size() > 4 ? col[4] : std::to_string(i);
            const std::string path = out_dir + "/out_" + lig_id + ".txt";
            if (!write_dump(path, col.size() > 3 ? col[3].c_str() : "", col.size() > 5 ? col[5].c_str() : "", cx,
                            docker)) {
                std::fprintf(stderr, "job %zu: cannot write %s\n", i, path.c_str());
                ++failed;
            }
            gd_complex_free(cxh);
            gd_raw_free(raw);
        }
    } catch (const std::exception &e) {
        std::fprintf(stderr, "%s\n", e.what());
        return EXIT_FAILURE;
    }
    return failed ? EXIT_FAILURE : EXIT_SUCCESS;
}
