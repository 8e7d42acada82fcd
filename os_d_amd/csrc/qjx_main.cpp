// qjx_main.cpp -- command-line drop-in for the reference QJX binary on the dimer
// STD/QJ path: reads config.txt and params.txt from the working directory (as
// the reference's main, main.cpp:3-192) and writes the same output files.
//
//   qjx [--config FILE] [--params FILE] [--precision fp64|fp32] [--steps-per-period S]
#include <cstdio>
#include <cstring>
#include <string>

#include "qjx.h"

int main(int argc, char **argv)
{
    std::string cfg = "config.txt", prm = "params.txt";
    int prec = QJX_FP64, steps = 0;
    for (int i = 1; i < argc; i++) {
        const char *a = argv[i];
        const bool has = i + 1 < argc;
        if (!std::strcmp(a, "--config") && has) cfg = argv[++i];
        else if (!std::strcmp(a, "--params") && has) prm = argv[++i];
        else if (!std::strcmp(a, "--steps-per-period") && has) steps = std::atoi(argv[++i]);
        else if (!std::strcmp(a, "--precision") && has) {
            const std::string p = argv[++i];
            if (p == "fp64") prec = QJX_FP64;
            else if (p == "fp32") prec = QJX_FP32;
            else if (p == "bf16x3") prec = QJX_BF16X3;
            else {
                std::fprintf(stderr, "unknown precision %s\n", p.c_str());
                return 2;
            }
        } else {
            std::fprintf(stderr, "usage: %s [--config FILE] [--params FILE] [--precision fp64|fp32|bf16x3] "
                                 "[--steps-per-period S]\n", argv[0]);
            return 2;
        }
    }
    const int rc = qjx_run_config(cfg.c_str(), prm.c_str(), prec, steps);
    if (rc != QJX_OK) {
        std::fprintf(stderr, "Error: %s\n", qjx_last_error());
        return 1;
    }
    return 0;
}
