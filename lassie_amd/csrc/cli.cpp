// LASSIE command line on liblassie:
//   LASSIE input_folder output_folder [-v1|-v2|-v3] [-p double|float]
// Same input files, defaults, output files and exit codes as the reference
// (src/c++/simulator.cu:29-345, src/c++/header/Simulation.cuh:230-981,
// src/c++/header/Reader.h:13-140); the integration runs on the GPU through the
// C ABI in include/lassie.h.
#include <dirent.h>
#include <sys/stat.h>
#include <sys/types.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include <hip/hip_runtime.h>

#include "../../include/lassie.h"

using std::string;
using std::vector;

namespace {

const char *kRule = "***************************************************************\n";

// ---- Reader.h semantics ----------------------------------------------------
// getline per line, '\t'-separated tokens, strtod (trailing text ignored, an
// empty token reads as 0).
vector<double> read_by_line(const string &f)  // Reader::readByLine, Reader.h:38-51
{
    vector<double> v;
    std::ifstream in(f.c_str());
    string line;
    while (std::getline(in, line)) v.push_back(strtod(line.c_str(), nullptr));
    return v;
}

vector<double> read_by_tab(const string &f)  // Reader::readByTab, Reader.h:53-69
{
    vector<double> v;
    std::ifstream in(f.c_str());
    string line;
    while (std::getline(in, line)) {
        std::istringstream buf(line);
        for (string tok; std::getline(buf, tok, '\t');) v.push_back(strtod(tok.c_str(), nullptr));
    }
    return v;
}

vector<int> read_by_line_int(const string &f)  // Reader::readByLineshort, Reader.h:89-102
{
    vector<int> v;
    std::ifstream in(f.c_str());
    string line;
    while (std::getline(in, line)) v.push_back((int)strtod(line.c_str(), nullptr));
    return v;
}

vector<vector<short>> read_matrix(const string &f)  // Reader::readMatrix, Reader.h:104-124
{
    vector<vector<short>> m;
    std::ifstream in(f.c_str());
    string line;
    while (std::getline(in, line)) {
        vector<short> row;
        std::istringstream buf(line);
        for (string tok; std::getline(buf, tok, '\t');) row.push_back((short)(int)strtod(tok.c_str(), nullptr));
        m.push_back(row);
    }
    return m;
}

double read_single(const string &f)  // Reader::readSingleValue, Reader.h:126-134
{
    std::ifstream in(f.c_str());
    string line;
    std::getline(in, line);
    return strtod(line.c_str(), nullptr);
}

int read_type(const string &f)  // Reader::readType, Reader.h:23-36
{
    std::ifstream in(f.c_str());
    string line;
    std::getline(in, line);
    if (line == "deterministic" || line == "concentration") return 0;
    if (line == "stochastic" || line == "number") return 1;
    return 2;
}

bool is_dir(const string &p)
{
    struct stat sb;
    return stat(p.c_str(), &sb) == 0 && S_ISDIR(sb.st_mode);
}

[[noreturn]] void die(int code, const string &msg)
{
    std::cout << msg;
    std::cout.flush();
    exit(code);
}

// The reference's launch rule and its record in infoBlockThread
// (returnBlockThread / minThread, Simulation.cuh:1225-1277, called for the
// ODE, Jacobian and sampling launches at :583-598): `thread`
// work items over `sm` compute units, at most 256 per block, rounded up to a
// multiple of 32; appends (save 1) or starts (save 0) the file.  Returns
// {blocks, threads per block}.  This build launches its own shapes; the
// record keeps the reference's file and its verbose report.
struct BlockThread {
    int blocks, tpb;
};

BlockThread block_thread(int sm, int thread, const string &file, const string &type, int save)
{
    FILE *fd = fopen(file.c_str(), save == 0 ? "w" : "a");
    if (fd) fprintf(fd, save == 0 ? "%s \n" : "\n\n%s \n", type.c_str());
    if (fd) fprintf(fd, "Allocating %d threads over %d available streaming multiprocessors\n", thread, sm);
    const float threadF = (float)thread;
    const int c = (int)ceilf(threadF / (float)sm);
    const int putative = 256 < c ? 256 : c;
    if (fd) fprintf(fd, "Putative TPB: %d \n", putative);
    float used_blocks, final_tpb;
    if (putative % 32 == 0) {
        final_tpb = (float)putative;
        used_blocks = ceilf((float)thread / final_tpb);
    } else {
        if (fd) fprintf(fd, "Adapting to warp size\n");
        final_tpb = (float)((putative / 32 + 1) * 32);
        used_blocks = ceilf((float)thread / final_tpb);
    }
    BlockThread bt{(int)used_blocks, (int)final_tpb};
    if (fd) {
        fprintf(fd, "Final TPB: %d \n", bt.tpb);
        fprintf(fd, "Number of used blocks: %d \n", bt.blocks);
        fprintf(fd, "Total threads spawned: %d \n", (int)(final_tpb * used_blocks));
        fclose(fd);
    }
    return bt;
}

}  // namespace

// ---- symbolic ODEs and Jacobian (verbose >= 2 / >= 3) ------------------------
// printOdes (Simulation.cuh:83-147): one line per non-fed species; a fed
// species prints nothing.  Terms follow the net-stoichiometry entries
// (reaction ascending), each "+c*k[r]" (or "-c*k[r]") followed by one "*x[s]"
// per unit of reactant order.
void print_odes(int S, const vector<int32_t> &lptr, const vector<int32_t> &lsp, const vector<int32_t> &lexp,
                const vector<int32_t> &hptr, const vector<int32_t> &hrx, const vector<int32_t> &hcoef,
                const vector<double> &feed)
{
    for (int g = 0; g < S; g++) {
        if (feed[g] != 0) continue;
        string s;
        for (int e = hptr[g]; e < hptr[g + 1]; e++) {
            const int r = hrx[e], c = hcoef[e];
            string term = (c > 0 ? "+" : "") + std::to_string(c) + "*k[" + std::to_string(r) + "]";
            for (int j = lptr[r]; j < lptr[r + 1]; j++)
                for (int l = 0; l < lexp[j]; l++) term += "*x[" + std::to_string(lsp[j]) + "]";
            s += term;
        }
        if (s.empty())
            printf("dx[%d] = 0\n", g);
        else
            printf("dx[%d] = %s\n", g, s.c_str());
    }
}

// printJac / printJac_ (Simulation.cuh:149-228): every entry (i, g), row
// major.  A column g of a fed species is 0.  For a reactant g of order 1 the
// term is "+c*k[r]" times the other reactants; of order e > 1 it is
// "(c*e)*k[r]" (no explicit "+"), x[g] repeated e-1 times, then the other
// reactants.  Other reactants appear once each, whatever their order.
void print_jac(int S, const vector<int32_t> &lptr, const vector<int32_t> &lsp, const vector<int32_t> &lexp,
               const vector<int32_t> &hptr, const vector<int32_t> &hrx, const vector<int32_t> &hcoef,
               const vector<double> &feed)
{
    for (int i = 0; i < S; i++) {
        for (int g = 0; g < S; g++) {
            if (feed[g] != 0) {
                printf("dx[%d][%d] = 0\n", i, g);
                continue;
            }
            string s;
            for (int e = hptr[i]; e < hptr[i + 1]; e++) {
                const int r = hrx[e], c = hcoef[e];
                for (int j = lptr[r]; j < lptr[r + 1]; j++) {
                    if (lsp[j] != g) continue;
                    string term;
                    if (lexp[j] == 1) {
                        term = (c > 0 ? "+" : "") + std::to_string(c) + "*k[" + std::to_string(r) + "]";
                    } else if (lexp[j] > 1) {
                        term = std::to_string(c * lexp[j]) + "*k[" + std::to_string(r) + "]";
                        for (int l = 0; l < lexp[j] - 1; l++) term += "*x[" + std::to_string(g) + "]";
                    } else {
                        continue;
                    }
                    for (int jj = lptr[r]; jj < lptr[r + 1]; jj++)
                        if (lsp[jj] != g) term += "*x[" + std::to_string(lsp[jj]) + "]";
                    s += term;
                }
            }
            if (s.empty())
                printf("dx[%d][%d] = 0\n", i, g);
            else
                printf("dx[%d][%d] = %s\n", i, g, s.c_str());
        }
    }
}

int main(int argc, char *argv[])
{
    const auto t_start = std::chrono::steady_clock::now();
    if (argc == 1)
        die(-3, string("\nError: the first parameter must be the input folder\n") +
                    "\nError: the second parameter must be the output folder\n" + kRule + "\n");
    if (argc == 2) die(-2, string("\nEror: the second parameter must be the output folder\n") + kRule + "\n");

    // ---- input folder (simulator.cu:77-152, openDirectory 322-345) ----------
    const string path = argv[1];
    DIR *dir = opendir(path.c_str());
    if (!dir)
        die(-1, string("\n\n") + kRule + "\n" + "\nError: the input folder does not exist\n" + kRule + "\n");
    string A, B, cvec, tvec, M0, csv = "NA", feed = "NA", kind = "NA", atolv = "NA", be = "NA", niter = "NA",
                                    ntol = "NA", rkf = "NA", stiff = "NA", volf = "NA";
    int have[5] = {0, 0, 0, 0, 0};
    for (struct dirent *e; (e = readdir(dir)) != nullptr;) {
        const string nm = e->d_name;
        if (nm == "." || nm == ".." || nm == ".DS_Store") continue;
        const string full = path + "/" + nm;
        if (nm == "left_side") A = full, have[0] = 1;
        else if (nm == "right_side") B = full, have[1] = 1;
        else if (nm == "c_vector") cvec = full, have[2] = 1;
        else if (nm == "t_vector") tvec = full, have[3] = 1;
        else if (nm == "M_0") M0 = full, have[4] = 1;
        else if (nm == "cs_vector") csv = full;
        else if (nm == "M_feed") feed = full;
        else if (nm == "modelkind") kind = full;
        else if (nm == "atol_vector") atolv = full;
        else if (nm == "be_step") be = full;
        else if (nm == "newton_iter") niter = full;
        else if (nm == "newton_tol") ntol = full;
        else if (nm == "rkf_step") rkf = full;
        else if (nm == "stiffness_tol") stiff = full;
        else if (nm == "volume") volf = full;
    }
    closedir(dir);
    const char *missing[5] = {"left side matrix", "right side matrix", "kinetic constants vector",
                              "vector of time instants", "vector of initial conditions"};
    bool to_exit = false;
    for (int i = 0; i < 5; i++)
        if (!have[i]) {
            std::cout << "\n\n" << kRule << "\n\nError: there is not " << missing[i] << "\n" << kRule << "\n";
            to_exit = true;
        }
    if (to_exit) exit(-4);

    // ---- output folder (simulator.cu:197-237) --------------------------------
    const string folder = argv[2];
    if (folder[0] == '-' || folder[0] == ' ')
        die(-2, string("\n\n") + kRule + "\n" + "\nError: the second parameter must be the output folder\n" + kRule +
                    "\n");
    if (!is_dir(folder)) mkdir(folder.c_str(), 0777);
    const string outdir = folder + "/output";
    if (!is_dir(outdir)) mkdir(outdir.c_str(), 0777);

    // ---- flags (simulator.cu:239-280) -----------------------------------------
    int verbose = 0;
    string prec = "double";
    for (int i = 3; i < argc; i++) {
        const string a = argv[i];
        if (a == "-v1") verbose = 1;
        if (a == "-v2") verbose = 2;
        if (a == "-v3") verbose = 3;
        if (a == "-p") {
            if (i == argc - 1)
                die(-5, string("\nError: you must insert the precison: double or float\n") + kRule + "\n");
            const string p = argv[i + 1];
            if (p != "double" && p != "float")
                die(-6, string("\nError: the parameter must be double or float\n") + kRule + "\n");
            prec = p;
        }
    }
    std::cout << "\n" << kRule << kRule << "\nLASSIE: A large-scale simulator of mass-action kinetics models\n";

    // ---- read the model (Simulation.cuh:250-391) ------------------------------
    vector<vector<short>> L = read_matrix(A), Rm = read_matrix(B);
    vector<double> k = read_by_line(cvec), t = read_by_line(tvec), x0 = read_by_tab(M0);
    const int R = (int)L.size();
    const int S = R > 0 ? (int)L[0].size() : 0;
    if (R < 1 || S < 1 || (int)Rm.size() != R || (int)k.size() < R || (int)x0.size() < S || t.empty())
        die(-4, string("\nError: inconsistent model files\n") + kRule + "\n");
    for (int r = 0; r < R; r++)
        if ((int)L[r].size() != S || (int)Rm[r].size() != S)
            die(-4, string("\nError: left_side and right_side rows must all have the same length\n") + kRule + "\n");
    const double dt_ei = be != "NA" ? read_single(be) : 0.01;
    const double dt_rkf = rkf != "NA" ? read_single(rkf) : 1e-3;
    const int iterN = niter != "NA" ? (int)read_single(niter) : 10000;
    const double tollN = ntol != "NA" ? read_single(ntol) : 1e-6;
    const double isStiff = stiff != "NA" ? read_single(stiff) : 1e-6;
    vector<double> atol = atolv != "NA" ? read_by_line(atolv) : vector<double>(S, 1e-12);
    vector<int> cs;
    if (csv != "NA")
        cs = read_by_line_int(csv);
    else
        for (int i = 0; i < S; i++) cs.push_back(i);
    vector<double> mfeed = feed != "NA" ? read_by_tab(feed) : vector<double>(S, 0.0);
    const int modelkind = kind != "NA" ? read_type(kind) : 0;
    atol.resize(S, 1e-12);
    mfeed.resize(S, 0.0);
    for (int c : cs)
        if (c < 0 || c >= S) die(-4, string("\nError: cs_vector entry out of range\n") + kRule + "\n");
    if (verbose) {
        std::cout << "\n" << kRule << "PARAMETERS:\nverbose = true\n";
        std::cout << "dt Runge-Kutta-Fehlberg = " << dt_rkf << "\n";
        std::cout << "dt Backward Euler = " << dt_ei << "\n";
        std::cout << "Newton's iterations = " << iterN << "\n";
        std::cout << "Newton's tollerance = " << tollN << "\n";
        std::cout << "Stiffness tollerance = " << isStiff << "\n";
        std::cout << "Runge-Kutta-Fehlberg's tollerances = " << atol[0] << "\n";
        std::cout << "Float pointing precision = " << prec << "\n";
    }
    if (modelkind == 2)
        die(-7, string("\nError: the modelkind can be: deterministic/concentration or stochastic/number\n") + kRule +
                    "\n");
    double volume = 1;
    const double avo = 6.022e23;
    if (modelkind) {
        // conversionToDeterministic, Simulation.cuh:983-1031
        volume = read_single(volf);
        if (volume == 0)
            die(-8, string("\nError: the modelkind is stochastic/number, but there is not volume file\n") + kRule +
                        "\n");
        const double conv = volume * avo;
        if (verbose > 1)
            std::cout << "\n\n" << kRule << "Convertion initial conditions and feeds with volume: " << volume
                      << std::endl;
        for (int i = 0; i < S; i++) {
            x0[i] = x0[i] / conv;
            mfeed[i] = mfeed[i] / conv;
        }
        for (int r = 0; r < R; r++) {
            if (r == 0 && verbose) std::cout << "Convertion reaction rate constants with volume: " << volume << std::endl;
            int distinct = 0;
            bool homo = false;
            for (int s = 0; s < S; s++) {
                distinct += L[r][s] > 0;
                homo = homo || L[r][s] >= 2;
            }
            if (distinct >= 3) {
                std::cout << "There are one or more reactions with higher order than the second one\n";
                exit(-12);
            }
            if (homo)
                k[r] = (k[r] * conv) / 2.0;
            else if (distinct > 1)
                k[r] = k[r] * conv;
        }
    }

    // ---- compile and upload -----------------------------------------------------
    vector<int16_t> Ld((size_t)R * S), Rd((size_t)R * S);
    for (int r = 0; r < R; r++)
        for (int s = 0; s < S; s++) {
            Ld[(size_t)r * S + s] = L[r][s];
            Rd[(size_t)r * S + s] = Rm[r][s];
        }
    int32_t nl = 0, nh = 0;
    if (lassie_csr_from_dense(S, R, Ld.data(), Rd.data(), &nl, &nh, nullptr, nullptr, nullptr, nullptr, nullptr,
                              nullptr) != LASSIE_OK)
        die(-4, string("\nError: ") + lassie_last_error() + "\n");
    vector<int32_t> lptr(R + 1), lsp(std::max(nl, 1)), lexp(std::max(nl, 1)), hptr(S + 1), hrx(std::max(nh, 1)),
        hcoef(std::max(nh, 1));
    lassie_csr_from_dense(S, R, Ld.data(), Rd.data(), nullptr, nullptr, lptr.data(), lsp.data(), lexp.data(),
                          hptr.data(), hrx.data(), hcoef.data());

    if (verbose > 1) {
        std::cout << "\n***************************************************************";
        std::cout << "\nODEs:\n";
        print_odes(S, lptr, lsp, lexp, hptr, hrx, hcoef, mfeed);
        std::cout << "***************************************************************\n";
        std::cout << "\n";
    }
    if (verbose > 2) {
        std::cout << "\n***************************************************************";
        std::cout << "\nJacobian:\n";
        print_jac(S, lptr, lsp, lexp, hptr, hrx, hcoef, mfeed);
        std::cout << "***************************************************************\n";
        std::cout << "\n";
    }

    // SpeciesSampling (Simulation.cuh:451-465)
    {
        FILE *fd = fopen((outdir + "/SpeciesSampling").c_str(), "w");
        if (!fd) {
            perror("Error open output file cs_vector");
            exit(-9);
        }
        fprintf(fd, "Saving species' dynamics: \n");
        for (int c : cs) fprintf(fd, "%d; ", c);
        fprintf(fd, "\n");
        fclose(fd);
    }

    // launch records (Simulation.cuh:583-620)
    int sm = 0;
    {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, 0) == hipSuccess) sm = prop.multiProcessorCount;
    }
    if (sm < 1) {
        fprintf(stderr, "Error: no HIP device\n");
        exit(-13);
    }
    const string info_file = outdir + "/infoBlockThread";
    const BlockThread bt_ode = block_thread(sm, S, info_file, "ODE", 0);
    const BlockThread bt_jac = block_thread(sm, S * S, info_file, "JAC", 1);
    block_thread(sm, (int)cs.size(), info_file, "DINAM-SUBSPEC", 1);
    if (verbose > 1) {
        std::cout << kRule;
        std::cout << "THREAD INFORMATIONS:\n";
        std::cout << "Number of blocks: " << bt_ode.blocks << std::endl;
        std::cout << "Number of threads: " << bt_ode.tpb << std::endl;
        std::cout << "Number of Jacobian blocks: " << bt_jac.blocks << std::endl;
        std::cout << "Number of Jacobian threads: " << bt_jac.tpb << std::endl;
        // print(), Simulation.cuh:1135-1184
        std::cout << "\n" << "***************************************************************";
        std::cout << "\nThe model has " << S << " species and " << R << " reactions" << "\n";
        if (!modelkind) {
            std::cout << "\n" << kRule << "modelkind = deterministic" << "\n";
        } else {
            std::cout << "\n" << kRule << "modelkind = stochastic" << "\n";
            std::cout << "\n" << kRule << "volume = " << volume << "\n";
        }
        std::cout << "\n\n###############################################################";
        std::cout << "\nStart ode solver\n\n";
        std::cout << "****Start simulation***" << std::endl;
    }

    lassie_model_desc d;
    d.n_species = S;
    d.n_reactions = R;
    d.react_ptr = lptr.data();
    d.react_species = lsp.data();
    d.react_exp = lexp.data();
    d.net_ptr = hptr.data();
    d.net_reaction = hrx.data();
    d.net_coef = hcoef.data();
    d.k = k.data();
    d.x0 = x0.data();
    d.feed = mfeed.data();
    d.atol = atol.data();
    lassie_model *model = nullptr;
    if (lassie_model_create(&d, &model) != LASSIE_OK) {
        fprintf(stderr, "Error %s\n", lassie_last_error());
        exit(-13);
    }
    lassie_opts o;
    lassie_opts_default(&o);
    o.rkf_step = dt_rkf;
    o.be_step = dt_ei;
    o.newton_iter = iterN;
    o.newton_tol = tollN;
    o.stiffness_tol = isStiff;
    // -p float: Simulation<float> (simulator.cu:259-304)
    const bool fp32 = prec == "float";
    if (fp32) o.precision = LASSIE_PREC_FP32_LOOP;
    const int nt = (int)t.size(), ncs = (int)cs.size();
    vector<double> dinamic((size_t)nt * ncs);
    // ProgressSimulation holds one line per loop iteration.  The solve keeps
    // the first prog_cap values; a longer run is solved again with room for
    // every iteration (the solve is deterministic, so the second run repeats
    // the first one), and a run too long to hold is an error.
    // (LASSIE_PROGRESS_BUFFER sets the first buffer's lines)
    int64_t prog_cap = 1 << 22;
    if (const char *pb = getenv("LASSIE_PROGRESS_BUFFER")) prog_cap = std::max<int64_t>(1, atoll(pb));
    const int64_t prog_max = (int64_t)1 << 28;
    vector<double> progress(prog_cap);
    lassie_stats st;
    int rc = lassie_solve(model, &o, t.data(), nt, cs.data(), ncs, dinamic.data(), nullptr, progress.data(), prog_cap,
                          &st);
    if (rc == LASSIE_OK && st.loop_iters > prog_cap) {
        if (st.loop_iters > prog_max) {
            fprintf(stderr, "Error: %lld loop iterations exceed the %lld lines ProgressSimulation can hold\n",
                    (long long)st.loop_iters, (long long)prog_max);
            exit(-14);
        }
        prog_cap = st.loop_iters;
        progress.assign(prog_cap, 0.0);
        rc = lassie_solve(model, &o, t.data(), nt, cs.data(), ncs, dinamic.data(), nullptr, progress.data(), prog_cap,
                          &st);
    }
    lassie_model_destroy(model);
    if (rc == LASSIE_ERR_MAXITER) {
        fprintf(stderr, "Error: %s\n", lassie_last_error());
        exit(-14);
    }
    if (rc != LASSIE_OK) {
        fprintf(stderr, "Error %s\n", lassie_last_error());
        exit(-13);
    }

    if (verbose > 1) {
        // the quarter marks the loop prints as samples are stored (Simulation.cuh:679-687)
        for (int64_t i = 1; i <= st.samples; i++) {
            if (i == nt / 4)
                std::cout << "****1/4 of simulation***" << std::endl;
            else if (i == nt / 2)
                std::cout << "****1/2 of simulation***" << std::endl;
            else if (i == 3 * nt / 4)
                std::cout << "****3/4 of simulation***" << std::endl;
        }
        std::cout << "****End simulation***" << std::endl;
    }

    // ProgressSimulation (Simulation.cuh:660-672): t/t_max per loop iteration
    {
        FILE *fd = fopen((outdir + "/ProgressSimulation").c_str(), "w");
        if (!fd) {
            perror("Error open output file ProgressSimulation");
            exit(-11);
        }
        for (int64_t i = 0; i < st.loop_iters; i++) fprintf(fd, "%.4e\n", progress[i]);
        fclose(fd);
    }
    // Solution (Simulation.cuh:942-971)
    {
        FILE *fd = fopen((outdir + "/Solution").c_str(), "w");
        if (!fd) {
            perror("Error open output file Solution");
            exit(-11);
        }
        if (modelkind)
            for (double &v : dinamic) v = v * (volume * avo);
        for (int i = 0; i < nt; i++) {
            fprintf(fd, "%.4e\t", fp32 ? (double)(float)t[i] : t[i]);  // t_vector holds T values
            for (int j = 0; j < ncs; j++) fprintf(fd, "%.4e\t", dinamic[(size_t)i * ncs + j]);
            fprintf(fd, "\n");
        }
        fclose(fd);
    }
    if (verbose > 1) {
        std::cout << "\nEnd ode solver\n";
        std::cout << "###############################################################\n";
    }
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
    std::cout << "\n" << kRule;
    printf("Total running time: %f seconds\n", secs);
    std::cout << kRule << kRule << "\n";
    return 0;
}
