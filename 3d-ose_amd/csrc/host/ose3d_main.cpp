// ose3d -- host driver of the MI355X OSE time stepper: the reference's
// main() (ose_3d.cu:70-309) with its hard-coded parameters turned into
// SU-style key=value arguments and presets.
//
//   ose3d preset=ref [vfile=input/v3d_nx101_ny101_nz101.bin] out=output
//   ose3d preset=c1 nt=200 out=/tmp/run snap_every=50 [snap_format=q8 snap_region=interior snap_lanes=q]
//   ose3d preset=c4 checkpoint=/scratch/ck checkpoint_every=500 [restore=/scratch/ck]
//
// Prints one JSON line with the model set-up and the measured rate.
#include <sys/stat.h>

#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "ose_host.h"

int main(int argc, char** argv)
{
    std::string params;
    std::string out = "output", ckpt, restore;
    int snap_every = 0, ckpt_every = 0;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        if (a == "-h" || a == "--help") {
            std::printf("usage: ose3d [preset=ref|c0|c1|c2|c3|c4] [key=value ...] [out=dir]\n"
                        "             [snap_every=N snap_format=f32|bf16|q16|q8 snap_region=padded|interior\n"
                        "              snap_lanes=pq|p|q] [checkpoint=dir checkpoint_every=N] [restore=dir]\n");
            return 0;
        }
        if (a.rfind("out=", 0) == 0) out = a.substr(4);
        else if (a.rfind("snap_every=", 0) == 0) snap_every = std::atoi(a.c_str() + 11);
        else if (a.rfind("checkpoint_every=", 0) == 0) ckpt_every = std::atoi(a.c_str() + 17);
        else if (a.rfind("checkpoint=", 0) == 0) ckpt = a.substr(11);
        else if (a.rfind("restore=", 0) == 0) restore = a.substr(8);
        params += a + " ";
    }
    oseh_sim* sim = nullptr;
    int rc = oseh_sim_create(&sim, params.c_str());
    if (rc) {
        std::fprintf(stderr, "ose3d: bad parameters (status %d)\n", rc);
        return 2;
    }
    if ((rc = oseh_sim_setup(sim))) {
        std::fprintf(stderr, "ose3d: set-up failed: %s\n", oseh_sim_error(sim));
        oseh_sim_destroy(sim);
        return 1;
    }
    mkdir(out.c_str(), 0755);
    std::vector<char> info(1 << 16);
    oseh_sim_info(sim, info.data(), info.size());
    // nt from the info JSON is the run length
    const std::string inf(info.data());
    const size_t p = inf.find("\"nt\": ");
    const int nt = p == std::string::npos ? 0 : std::atoi(inf.c_str() + p + 6);
    const size_t q = inf.find("\"interior_cells\": ");
    const double cells = q == std::string::npos ? 0.0 : std::atof(inf.c_str() + q + 18);
    double total_ms = 0.0;
    int done = 0;
    if (!restore.empty()) {
        if ((rc = oseh_sim_restore(sim, restore.c_str()))) {
            std::fprintf(stderr, "ose3d: restore failed: %s\n", oseh_sim_error(sim));
            oseh_sim_destroy(sim);
            return 1;
        }
        int step = 0;
        ose_get_step(oseh_sim_ctx(sim), &step);
        done = step;
        std::fprintf(stderr, "ose3d: resumed at step %d\n", done);
    }
    const int first = done;
    if (!ckpt.empty()) mkdir(ckpt.c_str(), 0755);
    // snapshots stream from the library while the steps run
    if (snap_every > 0 && (rc = oseh_sim_stream_snapshots(sim, out.c_str(), snap_every))) {
        std::fprintf(stderr, "ose3d: snapshots: %s\n", oseh_sim_error(sim));
        oseh_sim_destroy(sim);
        return 1;
    }
    while (done < nt) {
        int chunk = nt - done;
        if (ckpt_every > 0) chunk = std::min(chunk, ckpt_every - done % ckpt_every);
        double ms = 0.0;
        if ((rc = oseh_sim_run(sim, chunk, &ms))) {
            std::fprintf(stderr, "ose3d: run failed: %s\n", oseh_sim_error(sim));
            oseh_sim_destroy(sim);
            return 1;
        }
        total_ms += ms;
        done += chunk;
        if (!ckpt.empty() && ckpt_every > 0 && done % ckpt_every == 0 && done < nt &&
            (rc = oseh_sim_checkpoint(sim, ckpt.c_str()))) {
            std::fprintf(stderr, "ose3d: checkpoint failed: %s\n", oseh_sim_error(sim));
            oseh_sim_destroy(sim);
            return 1;
        }
        std::fprintf(stderr, "ose3d: step %d/%d\n", done, nt);
    }
    if ((rc = oseh_sim_write(sim, out.c_str()))) {
        std::fprintf(stderr, "ose3d: write failed: %s\n", oseh_sim_error(sim));
        oseh_sim_destroy(sim);
        return 1;
    }
    std::printf("{\"setup\": %s, \"steps\": %d, \"first_step\": %d, \"ms\": %.3f, \"gcell_steps_per_s\": %.4f}\n",
                info.data(), done, first, total_ms,
                total_ms > 0 ? cells * (done - first) / (total_ms * 1e6) : 0.0);
    oseh_sim_destroy(sim);
    return 0;
}
