/*
 * vfem_cli -- command-line drop-in for the reference's solver program
 * (src/main.cpp:28-127): reads a voxel model and its boundary conditions in the
 * reference's text formats (src/inout.cpp:6-128), assembles the system, solves it
 * on the GPU through the C++ layer (include/vfem.hpp) and prints the same
 * statistics and the compliance f.u, then writes <model>_out.csv like the
 * reference's saveVector (src/inout.cpp:207-225).
 *
 *     vfem_cli <model>.dat [-n levels] [--skip k]
 *              [--precision fp64|fp32|bf16|mixed] [--cycle reference|damped]
 *              [--omega w] [--cheb-level l] [--tol t] [--no-save]
 *
 * Files, relative to the working directory as in the reference:
 *   <model>.dat          "X Y Z" then Z lines of X*Y comma-separated integers (x fastest);
 *                        a voxel is solid only where its value is 1 (src/FEM.cpp:75)
 *   <model>_Ke0.dat      24 lines of 24 comma-separated values (else data/Ke0.dat)
 *   <model>_fixed.dat    count, then one 1-based node id per line
 *   <model>_loading.dat  count, then "id<TAB>fx<TAB>fy<TAB>fz" (1-based id)
 * Options beyond the reference's (-n, --skip) select the solver plan; the
 * defaults reproduce the reference: FP64, its additive cycle and its compiled
 * tolerance 1e-6 (include/FEMDefines.hpp:7).
 */
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <set>
#include <sstream>
#include <string>
#include <vector>

#include "vfem.hpp"

namespace {

struct Args {
    std::string model;
    int levels = 3;
    int skip = 0;
    int precision = VFEM_FP64;
    int cycle = VFEM_CYCLE_REFERENCE;
    double omega = -1;
    int cheb_level = 0;  // damped cycle: Chebyshev smoothing on levels >= cheb_level (0: damped Jacobi)
    double tol = -1;
    bool save = true;
};

[[noreturn]] void fail(const std::string &msg)
{
    std::cerr << "vfem_cli: " << msg << std::endl;
    std::exit(EXIT_FAILURE);
}

bool open_file(std::ifstream &in, const std::string &path)
{
    in.open(path);
    return (bool)in;
}

// <model>.dat: dimensions, then one comma-separated line of X*Y values per z layer
std::vector<int> read_voxels(const std::string &model, int &nx, int &ny, int &nz)
{
    std::ifstream in;
    if (!open_file(in, model + ".dat")) fail("cannot open " + model + ".dat");
    if (!(in >> nx >> ny >> nz) || nx <= 0 || ny <= 0 || nz <= 0) fail("bad dimensions in " + model + ".dat");
    std::vector<int> v((size_t)nx * ny * nz);
    std::string tok;
    std::getline(in, tok);  // rest of the dimension line
    size_t k = 0;
    std::string line;
    for (int z = 0; z < nz; z++) {
        if (!std::getline(in, line)) fail("missing voxel layer " + std::to_string(z) + " in " + model + ".dat");
        std::stringstream ss(line);
        for (int j = 0; j < nx * ny; j++) {
            if (!std::getline(ss, tok, ',')) fail("short voxel layer " + std::to_string(z) + " in " + model + ".dat");
            v[k++] = std::atoi(tok.c_str());  // passed through: only 1 is solid (src/FEM.cpp:75)
        }
    }
    return v;
}

void read_ke(const std::string &model, double ke[24][24])
{
    std::ifstream in;
    if (!open_file(in, model + "_Ke0.dat") && !open_file(in, "data/Ke0.dat"))
        fail("cannot open " + model + "_Ke0.dat or data/Ke0.dat");
    std::string line, tok;
    for (int i = 0; i < 24; i++) {
        if (!std::getline(in, line)) fail("Ke: missing row " + std::to_string(i));
        std::stringstream ss(line);
        for (int j = 0; j < 24; j++) {
            if (!std::getline(ss, tok, ',')) fail("Ke: short row " + std::to_string(i));
            ke[i][j] = std::strtod(tok.c_str(), nullptr);
        }
    }
}

void read_conditions(const std::string &model, std::set<uint64_t> &fixed,
                     std::map<uint64_t, vfem::Vec3<double>> &loads)
{
    std::ifstream in;
    if (!open_file(in, model + "_fixed.dat")) fail("cannot open " + model + "_fixed.dat");
    long long n = 0;
    if (!(in >> n) || n < 0) fail("bad count in " + model + "_fixed.dat");
    for (long long i = 0; i < n; i++) {
        unsigned long long id;
        if (!(in >> id) || id == 0) fail("bad fixed node id in " + model + "_fixed.dat");
        fixed.insert(id - 1);  // 1-based in the file
    }
    std::ifstream ld;
    if (!open_file(ld, model + "_loading.dat")) fail("cannot open " + model + "_loading.dat");
    if (!(ld >> n) || n < 0) fail("bad count in " + model + "_loading.dat");
    for (long long i = 0; i < n; i++) {
        unsigned long long id;
        double fx, fy, fz;
        if (!(ld >> id >> fx >> fy >> fz) || id == 0) fail("bad load line in " + model + "_loading.dat");
        loads[id - 1] = vfem::Vec3<double>{fx, fy, fz};
    }
}

Args parse(int argc, char **argv)
{
    if (argc < 2) fail("an input file is necessary: vfem_cli <model>.dat [-n levels] [--skip k] ...");
    Args a;
    const std::string file = argv[1];
    a.model = file.substr(0, file.find('.', 1));  // the reference's model name rule (main.cpp:38)
    for (int i = 2; i < argc; i++) {
        const std::string k = argv[i];
        auto val = [&]() -> std::string {
            if (i + 1 >= argc) fail("missing value after " + k);
            return argv[++i];
        };
        if (k == "-n") a.levels = std::atoi(val().c_str());
        else if (k == "--skip") a.skip = std::atoi(val().c_str());
        else if (k == "--precision") {
            const std::string p = val();
            if (p == "fp64") a.precision = VFEM_FP64;
            else if (p == "fp32") a.precision = VFEM_FP32;
            else if (p == "bf16") a.precision = VFEM_BF16;
            else if (p == "mixed") a.precision = VFEM_MIXED;
            else fail("unknown precision " + p);
        } else if (k == "--cycle") {
            const std::string c = val();
            if (c == "reference") a.cycle = VFEM_CYCLE_REFERENCE;
            else if (c == "damped") a.cycle = VFEM_CYCLE_DAMPED;
            else fail("unknown cycle " + c);
        } else if (k == "--omega") a.omega = std::strtod(val().c_str(), nullptr);
        else if (k == "--cheb-level") a.cheb_level = std::atoi(val().c_str());
        else if (k == "--tol") a.tol = std::strtod(val().c_str(), nullptr);
        else if (k == "--no-save") a.save = false;
        else fail("unknown option " + k);
    }
    return a;
}

// the reference's saveVector: one line, comma-separated, std::to_string values
void save_vector(const std::vector<double> &u, const std::string &model)
{
    std::ofstream out(model + "_out.csv");
    if (!out) fail("cannot write " + model + "_out.csv");
    for (size_t i = 0; i < u.size(); i++) out << (i ? "," : "") << std::to_string(u[i]);
    out << "\n";
}

}  // namespace

int main(int argc, char **argv)
{
    const Args a = parse(argc, argv);
    const auto t0 = std::chrono::steady_clock::now();
    int nx, ny, nz;
    std::vector<int> voxels = read_voxels(a.model, nx, ny, nz);
    std::cout << "Number of levels for Multigrid: " << a.levels;
    if (a.skip) std::cout << " (+ " << a.skip << " jump(s))";
    std::cout << std::endl;
    double ke[24][24];
    read_ke(a.model, ke);
    std::set<uint64_t> fixed;
    std::map<uint64_t, vfem::Vec3<double>> loads;
    read_conditions(a.model, fixed, loads);
    try {
        vfem::MatrixFreeSparse<double> K =
            vfem::assembleSystemMatrix<double>(voxels.data(), vfem::Vec3i{nx, ny, nz}, ke, fixed, a.levels, a.skip);
        const int64_t ndof = K.rows();
        long long solid = 0;
        for (int v : voxels) solid += (v == 1);
        std::cout << "System statistics:" << std::endl;
        std::cout << "\tVoxel model resolution: " << nx << " x " << ny << " x " << nz << std::endl;
        std::cout << "\tElements: " << solid << std::endl;
        std::cout << "\tDegrees of freedom: " << ndof << std::endl;
        std::vector<double> f(ndof, 0.0), u;
        vfem::applyBoundaryConditions(f, loads, fixed);
        const auto t1 = std::chrono::steady_clock::now();
        std::cout << "Preparation took:    " << std::chrono::duration<double>(t1 - t0).count() << " s" << std::endl;
        vfem_options o = K.options();
        o.precision = a.precision;
        o.cycle = a.cycle;
        if (a.omega > 0) o.omega = a.omega;
        o.cheb_level = a.cheb_level;
        if (a.tol >= 0) o.tol = a.tol;
        K.setOptions(o);
        vfem::solveWithHIP(K, f, u);
        const vfem_stats &st = K.lastStats();
        std::cout << "Solving took:        " << st.solve_ms / 1000.0 << " s (" << st.iterations
                  << " iterations, true relative residual " << st.rel_residual << ")" << std::endl;
        double compliance = 0;
        for (int64_t i = 0; i < ndof; i++) compliance += f[i] * u[i];
        std::cout << "Compliance:      " << compliance << std::endl;
        if (a.save) {
            std::cout << "Saving the solution" << std::endl;
            save_vector(u, a.model);
        }
        return st.converged ? EXIT_SUCCESS : 2;
    } catch (const vfem::Error &e) {
        fail(e.what());
    }
}
