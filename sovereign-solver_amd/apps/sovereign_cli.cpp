// sovereign-cli: the drop-in entry for `sovereign-cli solve --engine pdhg`.
// Mirrors the reference CLI's info / solve commands and the JSON fields of
// command_solve_pdhg (reference apps/sovereign_cli.cpp:196-236).  The other
// engines (simplex, ipm, race, crossover) are outside this build; asking for
// them exits with status 2 and says so.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "sovereign/backend.hpp"
#include "sovereign/io.hpp"
#include "sovereign/pdhg.hpp"

namespace {

std::string escape(const std::string& s) {
    std::string o;
    for (char c : s) {
        if (c == '"' || c == '\\') o += '\\';
        o += c;
    }
    return o;
}

int usage() {
    std::fprintf(stderr,
                 "usage: sovereign-cli info <model>\n"
                 "       sovereign-cli solve <model> [--engine pdhg] [--json] [--tolerance T]\n"
                 "                       [--max-iterations N] [--time-limit S] [--verbose]\n"
                 "                       [--backend host|hip] [--precision f64|f32]\n"
                 "                       [--termination reference|relative] [--primal-weight]\n");
    return 2;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 3) return usage();
    const std::string cmd = argv[1], path = argv[2];
    std::string engine = "pdhg";
    bool json = false;
    sov::PdhgOptions opt;
    sov::BackendKind kind = sov::default_backend_kind();
    for (int i = 3; i < argc; ++i) {
        const std::string a = argv[i];
        auto val = [&]() -> std::string {
            if (i + 1 >= argc) {
                std::fprintf(stderr, "%s needs a value\n", a.c_str());
                std::exit(2);
            }
            return argv[++i];
        };
        if (a == "--engine") engine = val();
        else if (a == "--json") json = true;
        else if (a == "--tolerance") opt.tolerance = std::strtod(val().c_str(), nullptr);
        else if (a == "--max-iterations") opt.max_iterations = std::strtoll(val().c_str(), nullptr, 10);
        else if (a == "--time-limit") opt.time_limit_seconds = std::strtod(val().c_str(), nullptr);
        else if (a == "--verbose") opt.verbose = true;
        else if (a == "--backend") {
            const std::string b = val();
            kind = b == "hip" ? sov::BackendKind::Hip : sov::BackendKind::Host;
        } else if (a == "--precision") opt.precision = val() == "f32" ? sov::Precision::F32 : sov::Precision::F64;
        else if (a == "--termination") opt.termination = val() == "relative" ? sov::PdhgTermination::RelativeKkt : sov::PdhgTermination::Reference;
        else if (a == "--primal-weight") opt.use_primal_weight = true;
        else return usage();
    }
    try {
        const sov::ReadResult r = sov::read_model(path);
        if (cmd == "info") {
            std::printf("%s\n%s\n", r.problem.summary().c_str(), sov::backend_report().c_str());
            return 0;
        }
        if (cmd != "solve") return usage();
        if (engine != "pdhg") {
            std::fprintf(stderr, "engine '%s' is not part of this build (PDHG only)\n", engine.c_str());
            return 2;
        }
        auto backend = sov::make_backend(kind);
        const sov::PdhgResult res = sov::Pdhg(opt).solve(r.problem, *backend);
        if (json) {
            std::printf("{\n  \"path\": \"%s\",\n  \"engine\": \"pdhg\",\n  \"backend\": \"%s\",\n", escape(path).c_str(),
                        sov::to_string(backend->kind()));
            std::printf("  \"status\": \"%s\",\n", sov::to_string(res.status));
            if (res.status == sov::PdhgStatus::Optimal) std::printf("  \"objective\": %.17g,\n", res.objective);
            else std::printf("  \"objective\": null,\n");
            std::printf("  \"iterations\": %lld,\n  \"restarts\": %lld,\n  \"seconds\": %.6f,\n", res.iterations,
                        res.restarts, res.seconds);
            std::printf("  \"primal_infeasibility\": %.17g,\n  \"dual_infeasibility\": %.17g,\n",
                        res.primal_infeasibility, res.dual_infeasibility);
            std::printf("  \"relative_kkt\": [%.6g, %.6g, %.6g],\n", res.relative_primal_residual,
                        res.relative_dual_residual, res.relative_gap);
            std::printf("  \"message\": \"%s\"\n}\n", escape(res.message).c_str());
        } else {
            std::printf("%s\n", r.problem.summary().c_str());
            std::printf("engine      : pdhg/%s (%s)\n", res.engine.c_str(), std::string(backend->device_description()).c_str());
            std::printf("status      : %s\n", sov::to_string(res.status));
            if (res.status == sov::PdhgStatus::Optimal) std::printf("objective   : %.12g\n", res.objective);
            std::printf("iterations  : %lld  (restarts: %lld)\n", res.iterations, res.restarts);
            std::printf("residuals   : primal %.3g, dual %.3g, relative kkt %.3g/%.3g/%.3g\n",
                        res.primal_infeasibility, res.dual_infeasibility, res.relative_primal_residual,
                        res.relative_dual_residual, res.relative_gap);
            std::printf("time        : %.3f s\n", res.seconds);
            if (!res.message.empty()) std::printf("note        : %s\n", res.message.c_str());
        }
        return res.status == sov::PdhgStatus::Optimal ? 0 : 1;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 2;
    }
}
