// Runs the integration guide's C++ adapters (NNPDevice::compute and the MD calculator CalcNNP) through
// libace_hip.so: usage adapter_test <model dump> <structure dump> <output>.
//   structure dump: n, the cell (9 numbers, column-major), then n lines "type x y z" (A);
//   output: "E1 e" (NNPDevice::compute), "F1 i fx fy fz" lines, then the calculator: "E2"/"F2" for the same
//   positions (forces added onto zeroed ones, as Engine::compute does) and "E3"/"F3" after atom 0 moved by
//   0.01 A along x.  Exit status 0 on success; a failure prints the library's message.
#include <cstdio>
#include <fstream>
#include <string>
#include "calc_nnp.hpp"
#include "nnp_device.hpp"

static Structure read_structure(const std::string& file) {
    std::ifstream in(file);
    if (!in) throw std::runtime_error("cannot open " + file);
    int n = 0;
    in >> n;
    Structure s;
    s.resize(n);
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) in >> s.R()(r, c);
    for (int i = 0; i < n; ++i) in >> s.type(i) >> s.posn(i)[0] >> s.posn(i)[1] >> s.posn(i)[2];
    if (!in) throw std::runtime_error("truncated structure dump " + file);
    return s;
}

static void write(FILE* f, const char* tag, double E, Structure& s) {
    std::fprintf(f, "E%s %.17g\n", tag, E);
    for (int i = 0; i < s.nAtoms(); ++i)
        std::fprintf(f, "F%s %d %.17g %.17g %.17g\n", tag, i, s.force(i)[0], s.force(i)[1], s.force(i)[2]);
}

static void zero_forces(Structure& s) {
    for (int i = 0; i < s.nAtoms(); ++i) s.force(i) = Eigen::Vector3d(0.0, 0.0, 0.0);
}

int main(int argc, char** argv) {
    if (argc != 4) {
        std::fprintf(stderr, "usage: %s <model dump> <structure dump> <output>\n", argv[0]);
        return 2;
    }
    try {
        const NNP nnp = read_dump(argv[1]);
        Structure s = read_structure(argv[2]);
        FILE* out = std::fopen(argv[3], "w");
        if (!out) throw std::runtime_error(std::string("cannot write ") + argv[3]);
        NNPDevice dev(nnp);
        const double E1 = dev.compute(s);
        write(out, "1", E1, s);
        CalcNNP calc(&dev, nnp.rc());
        calc.init(s);
        zero_forces(s);
        const double E2 = calc.compute(s);
        write(out, "2", E2, s);
        s.posn(0)[0] += 0.01;
        zero_forces(s);
        const double E3 = calc.compute(s);
        write(out, "3", E3, s);
        std::fclose(out);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "adapter_test: %s\n", e.what());
        return 1;
    }
    return 0;
}
