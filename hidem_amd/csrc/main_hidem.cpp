// hidem_run: the sample main (HiDEM/IceBlock/MainIceBlock.cpp,
// HiDEM/GlacierDem/MainDem.cpp) on the gfx950 accelerator.
//
//   hidem_run [--section.key=value ...] config.toml
//
// Reads [general] with the reference defaults and scaling, sets up the
// seeded FCC block from [lattice]/[hdm], and runs the reference loop:
// gridService, computePartForce, applyForce, linkUpdate (ts*dt > breakTime),
// finalizeTs.  [hdm] precision = "fp64" | "mixed" selects the precision form.
#include <chrono>
#include <cstdio>
#include <string>

#include "hidemlib_hip.h"

int main(int argc, char *argv[]) {
    try {
        hdm::ArgParser arguments(argc, argv);
        hdm::GeneralParams gp = hdm::GeneralParams::read(arguments);
        std::string prec = arguments.getArg<std::string>("hdm.precision").value_or("fp64");
        if (prec != "fp64" && prec != "mixed") throw hdm::TomlError("hdm.precision must be \"fp64\" or \"mixed\"");
        hdm::HiDEMLibHip hiDem(prec == "mixed" ? HDM_PREC_MIXED : HDM_PREC_FP64);
        hiDem.setup(arguments, gp.scale);

        std::printf("---------------------------------------\nUsing config\n\n%s", arguments.table.dump().c_str());
        std::printf("---------------------------------------\n");
        std::printf("particles %lld links %lld searchDistance %g interactDist %g precision %s\n",
                    (long long)hiDem.numParticles(), (long long)hiDem.numLinks(), hiDem.searchDistance,
                    hiDem.interactDist, prec.c_str());

        auto neighbourForceFunc = gp.neighForce();
        auto linkForceFunc = gp.linkForce();
        auto breakLinks = gp.linkBreak();
        hdm_integ_params applyForce = gp.integrator(arguments);

        auto iterStart = std::chrono::steady_clock::now();
        while (hiDem.run()) {
            hiDem.gridService();
            hiDem.computePartForce(neighbourForceFunc, linkForceFunc);
            hiDem.applyForce(applyForce);
            if (hiDem.getCurrentTimestep() * gp.dt > gp.breakTime) hiDem.linkUpdate(breakLinks);
            hiDem.finalizeTs();
        }
        hiDem.accel.sync();
        hiDem.finalize();
        double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - iterStart).count();
        int64_t broken = 0;
        hidem::hdm_check(hiDem.accel.ctx, hdm_count_broken(hiDem.accel.ctx, &broken), "hdm_count_broken");
        std::printf("Runtime %.3f s, %d steps, %.4e particle-steps/s, %lld rebuilds, %lld broken links\n", secs,
                    hiDem.getCurrentTimestep(), (double)hiDem.numParticles() * hiDem.getCurrentTimestep() / secs,
                    (long long)hiDem.rebuilds(), (long long)broken);
        return 0;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
}
