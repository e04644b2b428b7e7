// Text input formats of the distfind tool (cpp/distfind/cli/distfind_cli.cpp:100-188):
// whitespace-separated 0/1 matrices, and Pauli stabiliser strings split into the X-type
// and Z-type check matrices of a CSS code.
#include <cstdint>
#include <cstring>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/distfind_capi.h"

namespace qsa {
void set_last_error(const std::string& msg);   // capi.cpp
}

namespace {

int fail(const std::string& msg) {
    qsa::set_last_error(msg);
    return 2;
}

std::vector<std::string> lines_of(const char* text) {
    std::vector<std::string> out;
    std::istringstream in(text ? text : "");
    std::string line;
    while (std::getline(in, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        out.push_back(line);
    }
    return out;
}

}  // namespace

extern "C" {

// One matrix row per line, tokens "0" / "1" separated by whitespace, blank lines skipped
// (distfind_cli.cpp:100-129). out: rows x cols uint8, row-major, `cap` bytes.
int distfind_parse_matrix_text(const char* text, uint8_t* out, int64_t cap, int* rows, int* cols) {
    if (!text || !rows || !cols) return fail("null argument") - 1;
    *rows = 0;
    *cols = 0;
    std::vector<uint8_t> flat;
    int width = -1, r = 0;
    for (const std::string& line : lines_of(text)) {
        std::istringstream ls(line);
        std::string tok;
        int w = 0;
        while (ls >> tok) {
            if (tok != "0" && tok != "1") return fail("matrix contains non-binary token '" + tok + "'");
            flat.push_back(tok == "1");
            ++w;
        }
        if (w == 0) continue;
        if (width < 0) width = w;
        else if (w != width) return fail("ragged matrix (rows have differing widths)");
        ++r;
    }
    if (r == 0) return fail("matrix is empty");
    *rows = r;
    *cols = width;
    if ((int64_t)flat.size() > cap) return fail("output buffer too small");
    if (out) std::memcpy(out, flat.data(), flat.size());
    return 0;
}

// Pauli strings, one stabiliser per line; I . _ and space are identity; a blank line ends
// the input. Every stabiliser must be pure X or pure Z (CSS); all-identity rows are
// dropped (distfind_cli.cpp:133-188). hx / hz: row-major uint8, `cap` bytes each.
int distfind_parse_pauli_css(const char* text, uint8_t* hx, uint8_t* hz, int64_t cap, int* hx_rows, int* hz_rows,
                             int* n_out) {
    if (!text || !hx_rows || !hz_rows || !n_out) return fail("null argument") - 1;
    *hx_rows = *hz_rows = *n_out = 0;
    std::vector<std::string> lines;
    for (const std::string& l : lines_of(text)) {
        if (l.empty()) break;
        lines.push_back(l);
    }
    if (lines.empty()) return fail("no stabilisers (provide Pauli strings or 0/1 matrices)");
    const int n = (int)lines[0].size();
    for (const std::string& s : lines)
        if ((int)s.size() != n) return fail("stabilisers have differing lengths (all rows must span the same qubits)");
    std::vector<uint8_t> xs, zs;
    int nx = 0, nz = 0;
    for (size_t li = 0; li < lines.size(); ++li) {
        std::vector<uint8_t> x(n, 0), z(n, 0);
        bool hasx = false, hasz = false;
        for (int j = 0; j < n; ++j) {
            const char c = lines[li][j];
            if (c == 'I' || c == '.' || c == '_' || c == ' ') continue;
            if (c == 'X' || c == 'x') { x[j] = 1; hasx = true; continue; }
            if (c == 'Z' || c == 'z') { z[j] = 1; hasz = true; continue; }
            if (c == 'Y' || c == 'y')
                return fail("stabiliser " + std::to_string(li + 1) + " contains a Y -- only CSS codes are supported");
            return fail("stabiliser " + std::to_string(li + 1) + " has an unrecognised character '" + std::string(1, c) + "'");
        }
        if (hasx && hasz)
            return fail("stabiliser " + std::to_string(li + 1) + " mixes X and Z -- only CSS codes are supported");
        if (hasx) { xs.insert(xs.end(), x.begin(), x.end()); ++nx; }
        else if (hasz) { zs.insert(zs.end(), z.begin(), z.end()); ++nz; }
    }
    *hx_rows = nx;
    *hz_rows = nz;
    *n_out = n;
    if ((int64_t)xs.size() > cap || (int64_t)zs.size() > cap) return fail("output buffer too small");
    if (hx && !xs.empty()) std::memcpy(hx, xs.data(), xs.size());
    if (hz && !zs.empty()) std::memcpy(hz, zs.data(), zs.size());
    return 0;
}

}  // extern "C"
