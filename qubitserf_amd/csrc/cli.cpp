// distfind command-line tool over the gfx950 engine: the reference tool's input formats,
// flags and output (cpp/distfind/cli/distfind_cli.cpp), method "bz" on the device.
//
//   cat code.txt | distfind [--z | --x | --zx] [-v]
//   distfind --hx Hx.txt --hz Hz.txt [--z | --x | --zx]
//
// stdout holds just the number(s); an unproven result is noted on stderr. Methods and
// modes this engine does not provide (--cc, --mitm, --cpu, --subsystem, -o) exit with
// status 2 and a message, like any other usage error.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/distfind_capi.h"

namespace {

const char* kUsage =
    "Usage: distfind [OPTIONS]\n"
    "       cat code.txt | distfind [OPTIONS]\n"
    "       distfind --hx Hx.txt --hz Hz.txt [OPTIONS]\n"
    "\n"
    "Minimum distance of a CSS code by Brouwer-Zimmermann on a gfx950 GPU.\n"
    "\n"
    "Input (default): Pauli stabiliser strings on stdin, one per line (I . _ identity,\n"
    "X, Z); a blank line or EOF ends the input. Each must be pure X- or Z-type.\n"
    "Output:  default = minimum distance; --zx = \"<dZ> <dX>\"; --z / --x = one value\n"
    "Other:   --bz  --gpu  --threads N (accepted, unused)  --hx FILE --hz FILE  -v  -h\n";

[[noreturn]] void die(const std::string& msg) {
    std::cerr << "distfind: " << msg << "\n";
    std::exit(2);
}

std::string slurp(std::istream& in) {
    std::ostringstream ss;
    ss << in.rdbuf();
    return ss.str();
}

struct Mat {
    std::vector<uint8_t> a;
    int rows = 0, cols = 0;
};

Mat read_matrix(const std::string& path) {
    std::ifstream f(path);
    if (!f) die("cannot open file: " + path);
    const std::string text = slurp(f);
    Mat m;
    m.a.resize(text.size() + 1);
    if (distfind_parse_matrix_text(text.c_str(), m.a.data(), (int64_t)m.a.size(), &m.rows, &m.cols) != 0)
        die(std::string(distfind_last_error()) + " in " + path);
    m.a.resize((size_t)m.rows * m.cols);
    return m;
}

void note_unproven(const DistfindResult& r, const char* label) {
    if (!r.proven) std::cerr << "distfind: " << label << " d in [" << r.lower_bound << ", " << r.distance << "] (not proven)\n";
}

DistfindResult solve(const Mat& hx, const Mat& hz, char which, bool verbose) {
    DistfindResult r{};
    const int rc = distfind_css_distance(hx.a.data(), hx.rows, hx.cols, hz.a.data(), hz.rows, hz.cols, "bz", which, "gpu",
                                         0, verbose ? 1 : 0, &r);
    if (rc != 0) die(distfind_last_error());
    return r;
}

}  // namespace

int main(int argc, char** argv) {
    char which = 'M';
    bool zx = false, verbose = false;
    std::string hx_path, hz_path;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto value = [&](const std::string& flag) -> std::string {
            if (i + 1 >= argc) die(flag + " requires a value");
            return argv[++i];
        };
        if (a == "-h" || a == "--help") { std::cout << kUsage; return 0; }
        else if (a == "--bz" || a == "--gpu") {}
        else if (a == "--cc" || a == "--mitm") die(a.substr(2) + " is not provided by the gfx950 engine (it runs --bz)");
        else if (a == "--cpu") die("--cpu is served by the reference tool, not the gfx950 engine");
        else if (a == "--subsystem") die("--subsystem is not provided by the gfx950 engine");
        else if (a == "-o" || a == "--operator") die("operator weight is not provided by the gfx950 engine");
        else if (a == "--zx") zx = true;
        else if (a == "--z") which = 'Z';
        else if (a == "--x") which = 'X';
        else if (a == "--threads") {
            const std::string v = value(a);
            if (v.empty() || v.find_first_not_of("0123456789") != std::string::npos)
                die(a + " expects a non-negative integer, got '" + v + "'");
        }
        else if (a == "-v" || a == "--verbose") verbose = true;
        else if (a == "--hx") hx_path = value(a);
        else if (a == "--hz") hz_path = value(a);
        else die("unrecognised argument '" + a + "' (try --help)");
    }
    if (zx && which != 'M') die("--zx cannot be combined with --z or --x");
    if (hx_path.empty() != hz_path.empty()) die("--hx and --hz must be given together");

    Mat hx, hz;
    if (!hx_path.empty()) {
        hx = read_matrix(hx_path);
        hz = read_matrix(hz_path);
        if (hx.cols != hz.cols) die("Hx and Hz have differing column counts (different number of qubits)");
    } else {
        const std::string text = slurp(std::cin);
        std::vector<uint8_t> bx(text.size() + 1), bz(text.size() + 1);
        int n = 0;
        if (distfind_parse_pauli_css(text.c_str(), bx.data(), bz.data(), (int64_t)bx.size(), &hx.rows, &hz.rows, &n) != 0)
            die(distfind_last_error());
        hx.cols = hz.cols = n;
        hx.a.assign(bx.begin(), bx.begin() + (size_t)hx.rows * n);
        hz.a.assign(bz.begin(), bz.begin() + (size_t)hz.rows * n);
    }
    if (verbose)
        std::cerr << "distfind: method=bz backend=gpu qubits=" << hx.cols << " Hx_rows=" << hx.rows
                  << " Hz_rows=" << hz.rows << "\n";
    if (zx) {   // both sides, each to its own full proof
        const DistfindResult z = solve(hx, hz, 'Z', verbose), x = solve(hx, hz, 'X', verbose);
        note_unproven(z, "dZ:");
        note_unproven(x, "dX:");
        std::cout << z.distance << ' ' << x.distance << std::endl;
    } else {
        const DistfindResult r = solve(hx, hz, which, verbose);
        note_unproven(r, which == 'Z' ? "dZ:" : which == 'X' ? "dX:" : "d:");
        std::cout << r.distance << std::endl;
    }
    return 0;
}
