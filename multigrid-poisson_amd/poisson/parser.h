// Command-line settings (reference include/parser.h, src/parser.cpp:3-50).
#ifndef INPUT_PARSER
#define INPUT_PARSER

#include <cmath>
#include <iostream>
#include <string>

#include "definitions.h"
#include "settings.h"

namespace Poisson {
inline Settings parser(int argc, char* argv[]) {
    std::string str;
    Settings set;
    for (int i = 1; i < argc; i++) {
        str = argv[i];
        const bool has_value = i + 1 < argc;
        if (str == "-x" && has_value) {
            set.dims[0] = std::stoi(argv[++i]);
        } else if (str == "-y" && has_value) {
            set.dims[1] = std::stoi(argv[++i]);
        } else if (str == "-z" && has_value) {
            set.dims[2] = std::stoi(argv[++i]);
        } else if ((str == "-tolerance" || str == "-tol") && has_value) {
            set.tolerance = std::stod(argv[++i]);
        } else if (str == "-maxiter" && has_value) {
            set.maxiter = std::stoi(argv[++i]);
        } else if ((str == "-levels" || str == "-l") && has_value) {
            set.levels = std::stoi(argv[++i]);
        } else if (str == "-save" || str == "-s") {
            set.print_result = true;
        } else if (str == "-length" && has_value) {
            set.lengthx = std::stod(argv[++i]);
        } else if (str == "-stats" && has_value) {
            set.stats_file = argv[++i];
            set.write_final_stats = true;
        } else if (str == "-consistent") {  // addition: SolverMode CONSISTENT
            set.mode = CONSISTENT;
        } else {
            std::cerr << "Unknown input " << str << std::endl;
        }
    }
    set.h = set.lengthx / ((double_t)set.dims[0] - 1.0);  // isotropic spacing from x
    return set;
}
}  // namespace Poisson
#endif
