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
//   gd_dock <jobs.csv> [--out DIR] [--seed S] [--device D]
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
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

} // namespace

int main(int argc, char **argv)
{
    std::string fn, out_dir = ".";
    uint64_t seed = 2024;
    int device = 0;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "--out" && i + 1 < argc) out_dir = argv[++i];
        else if (a == "--seed" && i + 1 < argc) seed = std::strtoull(argv[++i], nullptr, 10);
        else if (a == "--device" && i + 1 < argc) device = std::atoi(argv[++i]);
        else if (fn.empty()) fn = a;
        else {
            std::fprintf(stderr, "%s <FILE(*.csv)> [--out DIR] [--seed S] [--device D]\n", argv[0]);
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
