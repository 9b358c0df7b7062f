// dda_fft: the command-line program of the DDA-FFT solve path on MI355X.
//
// A drop-in for the reference's dda_fft.out: the same options, defaults and
// validation (options.cpp:233-1155), the same state file with its resume
// semantics (main.cpp:214-262, 958-983) and the same out.dat layout
// (main.cpp:1085-1118).  The sample loop (main.cpp:774-996) is batched for
// the GPU: on a target whose occupancy does not change between samples, the
// Monte-Carlo samples of one group are solved together as right-hand sides of
// one call (two polarizations each, each with its own incidence and rotated
// directions); random targets regenerate their occupancy per sample in one
// context whose A-hat (a function of lattice and kd only) is built once.
// Samples follow the random stream of the reference's first worker, i.e. the
// results of `-j 1`; the reference's interactive prompts are not provided
// (missing required values are errors).
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>
#include <vector>

#include "../../include/ddafft.h"

namespace {

const double kPi = 3.1415926535897932384626433832795;
const double kUnset = std::nan("");
const int kUnsetInt = -2147483647;

struct Options {
    double theta_min = kUnset, theta_max = kUnset, theta_step = kUnset;
    std::string state_file, output_file;
    bool restart = false;
    int samples = kUnsetInt, orientation = kUnsetInt;
    double psi = kUnset, iota = kUnset, kappa = kUnset;
    int planes = kUnsetInt;
    int target = kUnsetInt;
    int grid[3] = {0, 0, 0};
    double size_parameter = kUnset, aspect_b = kUnset, aspect_c = kUnset, radius_ratio = kUnset;
    int levels = kUnsetInt, bites = kUnsetInt;
    double depth = kUnset, packing = kUnset, monomer_radius = kUnset, monomer_length_ratio = kUnset;
    std::string shape_file;
    std::vector<float> mat_re, mat_im;
    std::vector<int> seeds;
    int polarizability = kUnsetInt;
    double tolerance = kUnset;
    int max_steps = kUnsetInt, fft_mode = kUnsetInt, solver = kUnsetInt;
    double tau = kUnset;
    int threads = kUnsetInt;
    bool has_seed = false;
    unsigned long long seed = 0;
    int group = 8;   // samples per batched solve (this program's own option)
    bool reference_sums = false;   // observables summed as the reference sums them (own option)
    bool compact_ahat = false;     // A-hat kept as one octant, Nyquist planes of A(r) zeroed (own option)
};

struct Name {
    const char* name;
    int code;
};
const Name kTargets[] = {{"sphere", 1},        {"ellipsoid", 2},      {"parallelepiped", 3},
                         {"random-sphere", 4}, {"random-ellipsoid", 5}, {"anthrax", 6},
                         {"hedgehog", 7},      {"dead-anthrax", 8},   {"random-parallelepiped", 9},
                         {"nibble-ellipsoid", 10}, {"file", 11},      {"bisphere", 12},
                         {"hierarchical", 13}, {"sphere-spec", 14},   {nullptr, 0}};
const Name kOrient[] = {{"fixed", 0}, {"mc", 1}, {"monte-carlo", 1}, {nullptr, 0}};
const Name kPolar[] = {{"cm", 1}, {"cmrr", 2}, {"cmnr", 3}, {"ildr", 4}, {"ldr", 5}, {nullptr, 0}};
const Name kFft[] = {{"ct", 0}, {"cooley-tukey", 0}, {"rb", 1}, {"rader-brenner", 1}, {nullptr, 0}};
const Name kSolver[] = {{"cg", 0}, {"aa", 1}, {"bicgstab", 2}, {"simple", 3}, {nullptr, 0}};

bool parse_int(const std::string& s, int* v) {
    char* end;
    errno = 0;
    const long x = std::strtol(s.c_str(), &end, 10);
    if (end == s.c_str() || *end || errno) return false;
    *v = (int)x;
    return true;
}
bool parse_dbl(const std::string& s, double* v) {
    char* end;
    errno = 0;
    const double x = std::strtod(s.c_str(), &end);
    if (end == s.c_str() || *end || errno) return false;
    *v = x;
    return true;
}
// a number or one of the names (options.cpp:95-122)
bool parse_code(const Name* names, const std::string& s, int* v) {
    if (parse_int(s, v)) return true;
    for (int i = 0; names[i].name; ++i)
        if (s == names[i].name) {
            *v = names[i].code;
            return true;
        }
    return false;
}
bool split(const std::string& s, std::vector<std::string>* out) {
    out->clear();
    size_t a = 0;
    while (a <= s.size()) {
        const size_t b = s.find(',', a);
        const std::string t = s.substr(a, b == std::string::npos ? std::string::npos : b - a);
        if (t.empty()) return false;
        out->push_back(t);
        if (b == std::string::npos) break;
        a = b + 1;
    }
    return !out->empty();
}
bool looks_numeric(const char* s) {
    if (!s || !*s) return false;
    char* end;
    errno = 0;
    (void)std::strtod(s, &end);
    return end != s && *end == 0 && errno == 0;
}

void usage(const char* prog) {
    std::printf("\nUsage: %s [options]\n"
                "\n DDA-FFT light scattering on one MI355X (drop-in for the reference program).\n"
                "\nRun and output:\n"
                "  -o, --output <file>              text result file (default out.dat)\n"
                "  -D, --state <file>               binary state file, resumed if it exists (default out.dst)\n"
                "  --restart                        delete the state file and start from scratch\n"
                "  -n, --samples <N>                number of samples/orientations (default 1)\n"
                "  -a, --theta-min <deg>            lower limit of the phase angles (default 0)\n"
                "  -b, --theta-max <deg>            upper limit of the phase angles (default 180)\n"
                "  -s, --theta-step <deg>           phase-angle increment (default 10)\n"
                "  -O, --orientation <mode>         fixed|mc (default fixed)\n"
                "  --psi, --iota, --kappa <deg>     rotations, fixed orientation (default 0)\n"
                "  -p, --planes <N>                 number of scattering planes (default 1)\n"
                "\nTarget:\n"
                "  -t, --target <name|1..14>        target type\n"
                "  -g, --grid <i[,j,l]>             lattice dimensions, powers of two\n"
                "  -x, --size-parameter <x>         size parameter of the target\n"
                "  --aspect <b/a,c/a>  --radius-ratio <R1/R2>  --levels <N>  --bites <N>  --depth <r>\n"
                "  --packing <rho>  --monomer-radius <r>  --monomer-length-ratio <l/r>\n"
                "  -f, --shape-file <file>          file with the target shape, target 11\n"
                "  -m, --material <mre,mi>          refractive index of one material; repeat for a mixture\n"
                "  --seeds <seed[,seed...]>         random seeds, one per material\n"
                "\nNumerics:\n"
                "  -P, --polarizability <model>     1..5 or cm,cmrr,cmnr,ildr,ldr (default 1)\n"
                "  -e, --tolerance <err>            convergence threshold (default 1e-3)\n"
                "  -S, --max-steps <N>              maximum number of iterations (default 20)\n"
                "  -F, --fft <method>               0|1 or ct,rb (default 0; both map to the GPU transform)\n"
                "  -M, --solver <method>            0..3 or cg,aa,bicgstab,simple (default 0)\n"
                "  --tau <factor>                   weight factor, required by solver 3\n"
                "  -j, --threads <N>                accepted for compatibility (samples follow -j 1)\n"
                "  --seed <u64>                     random seed for reproducible runs\n"
                "  --group <N>                      samples solved together on the GPU, 1..32 (default 8)\n"
                "  --reference-sums                 sum Cext, Cabs and the far field sequentially in float, in the\n"
                "                                   reference's order (default: FP64 sums of the same terms)\n"
                "  --compact-ahat                   keep one octant of A-hat, the Nyquist planes of A(r) zeroed:\n"
                "                                   1/8 of its memory, results within float rounding of the\n"
                "                                   default (which is bit-identical to the reference's)\n",
                prog);
}

// Returns 0 ok, 1 help/version (exit 0), 2 error.
int parse(int argc, char** argv, Options* o) {
    // fold the values of list options into one token (options.cpp:186-276)
    struct ListOpt {
        const char* tok;
        const char* emit;
        int lo, hi;
    };
    const ListOpt lists[] = {{"-m", "--material", 2, 2}, {"--material", "--material", 2, 2},
                             {"-g", "--grid", 1, 3},     {"--grid", "--grid", 1, 3},
                             {"--aspect", "--aspect", 2, 2}, {"--seeds", "--seeds", 1, 64}};
    std::vector<std::string> args;
    for (int i = 1; i < argc;) {
        const ListOpt* lo = nullptr;
        for (const auto& l : lists)
            if (std::strcmp(argv[i], l.tok) == 0) lo = &l;
        if (!lo) {
            args.push_back(argv[i++]);
            continue;
        }
        std::string vals;
        int taken = 0;
        for (int j = i + 1; j < argc && taken < lo->hi; ++j) {
            if (!looks_numeric(argv[j])) break;
            if (taken) vals += ',';
            vals += argv[j];
            ++taken;
        }
        if (taken >= lo->lo) {
            args.push_back(std::string(lo->emit) + "=" + vals);
            i += taken + 1;
        } else {
            args.push_back(argv[i++]);
        }
    }
    std::vector<std::string> errors;
    auto need = [&](size_t& k, const std::string& key, std::string* val, const std::string& inl) -> bool {
        if (!inl.empty()) {
            *val = inl;
            return true;
        }
        if (k + 1 < args.size()) {
            *val = args[++k];
            return true;
        }
        errors.push_back("option " + key + " requires an argument");
        return false;
    };
    for (size_t k = 0; k < args.size(); ++k) {
        std::string a = args[k], key = a, inl;
        const size_t eq = a.find('=');
        if (a.rfind("--", 0) == 0 && eq != std::string::npos) {
            key = a.substr(0, eq);
            inl = a.substr(eq + 1);
        }
        std::string v;
        int iv;
        double dv;
        if (key == "-h" || key == "--help") {
            usage(argv[0]);
            return 1;
        } else if (key == "-V" || key == "--version") {
            std::printf("\n%s 1.0 (MI355X)\n", argv[0]);
            return 1;
        } else if (key == "-i" || key == "--interactive") {
            errors.push_back("--interactive: prompts are not provided; give every value as an option");
        } else if (key == "--restart") {
            o->restart = true;
        } else if (key == "--reference-sums") {
            o->reference_sums = true;
        } else if (key == "--compact-ahat") {
            o->compact_ahat = true;
        } else if (key == "-o" || key == "--output") {
            if (need(k, key, &v, inl)) o->output_file = v;
        } else if (key == "-D" || key == "--state") {
            if (need(k, key, &v, inl)) o->state_file = v;
        } else if (key == "-f" || key == "--shape-file") {
            if (need(k, key, &v, inl)) o->shape_file = v;
        } else if (key == "-n" || key == "--samples" || key == "-p" || key == "--planes" || key == "--levels" ||
                   key == "--bites" || key == "-S" || key == "--max-steps" || key == "-j" || key == "--threads" ||
                   key == "--group") {
            if (!need(k, key, &v, inl)) continue;
            if (!parse_int(v, &iv)) {
                errors.push_back(key + ": expected an integer, got \"" + v + "\"");
                continue;
            }
            if (key == "-n" || key == "--samples") o->samples = iv;
            if (key == "-p" || key == "--planes") o->planes = iv;
            if (key == "--levels") o->levels = iv;
            if (key == "--bites") o->bites = iv;
            if (key == "-S" || key == "--max-steps") o->max_steps = iv;
            if (key == "-j" || key == "--threads") o->threads = iv;
            if (key == "--group") o->group = iv;
        } else if (key == "-a" || key == "--theta-min" || key == "-b" || key == "--theta-max" || key == "-s" ||
                   key == "--theta-step" || key == "--psi" || key == "--iota" || key == "--kappa" || key == "-x" ||
                   key == "--size-parameter" || key == "--radius-ratio" || key == "--depth" || key == "--packing" ||
                   key == "--monomer-radius" || key == "--monomer-length-ratio" || key == "-e" ||
                   key == "--tolerance" || key == "--tau") {
            if (!need(k, key, &v, inl)) continue;
            if (!parse_dbl(v, &dv)) {
                errors.push_back(key + ": expected a number, got \"" + v + "\"");
                continue;
            }
            if (key == "-a" || key == "--theta-min") o->theta_min = dv;
            if (key == "-b" || key == "--theta-max") o->theta_max = dv;
            if (key == "-s" || key == "--theta-step") o->theta_step = dv;
            if (key == "--psi") o->psi = dv;
            if (key == "--iota") o->iota = dv;
            if (key == "--kappa") o->kappa = dv;
            if (key == "-x" || key == "--size-parameter") o->size_parameter = dv;
            if (key == "--radius-ratio") o->radius_ratio = dv;
            if (key == "--depth") o->depth = dv;
            if (key == "--packing") o->packing = dv;
            if (key == "--monomer-radius") o->monomer_radius = dv;
            if (key == "--monomer-length-ratio") o->monomer_length_ratio = dv;
            if (key == "-e" || key == "--tolerance") o->tolerance = dv;
            if (key == "--tau") o->tau = dv;
        } else if (key == "-O" || key == "--orientation") {
            if (!need(k, key, &v, inl)) continue;
            if (parse_code(kOrient, v, &iv) && (iv == 0 || iv == 1))
                o->orientation = iv;
            else
                errors.push_back("--orientation: expected fixed or mc, got \"" + v + "\"");
        } else if (key == "-t" || key == "--target") {
            if (!need(k, key, &v, inl)) continue;
            if (parse_code(kTargets, v, &iv) && iv >= 1 && iv <= 14)
                o->target = iv;
            else
                errors.push_back("--target: unknown target \"" + v + "\"");
        } else if (key == "-P" || key == "--polarizability") {
            if (!need(k, key, &v, inl)) continue;
            if (parse_code(kPolar, v, &iv) && iv >= 1 && iv <= 5)
                o->polarizability = iv;
            else
                errors.push_back("--polarizability: expected 1..5 or a name, got \"" + v + "\"");
        } else if (key == "-F" || key == "--fft") {
            if (!need(k, key, &v, inl)) continue;
            if (parse_code(kFft, v, &iv) && (iv == 0 || iv == 1))
                o->fft_mode = iv;
            else
                errors.push_back("--fft: expected 0|1 or ct|rb, got \"" + v + "\"");
        } else if (key == "-M" || key == "--solver") {
            if (!need(k, key, &v, inl)) continue;
            if (parse_code(kSolver, v, &iv) && iv >= 0 && iv <= 3)
                o->solver = iv;
            else
                errors.push_back("--solver: expected 0..3 or a name, got \"" + v + "\"");
        } else if (key == "--seed") {
            if (!need(k, key, &v, inl)) continue;
            char* end;
            errno = 0;
            const unsigned long long s = std::strtoull(v.c_str(), &end, 10);
            if (end == v.c_str() || *end || errno)
                errors.push_back("--seed: expected an unsigned integer, got \"" + v + "\"");
            else {
                o->has_seed = true;
                o->seed = s;
            }
        } else if (key == "--grid" || key == "--aspect" || key == "--material" || key == "--seeds") {
            if (!need(k, key, &v, inl)) continue;
            std::vector<std::string> t;
            if (!split(v, &t)) {
                errors.push_back(key + ": bad list \"" + v + "\"");
                continue;
            }
            if (key == "--grid") {
                std::vector<int> g;
                for (auto& x : t)
                    if (parse_int(x, &iv)) g.push_back(iv);
                if (g.size() != t.size())
                    errors.push_back("--grid: expected integers, got \"" + v + "\"");
                else if (g.size() == 1)
                    o->grid[0] = o->grid[1] = o->grid[2] = g[0];
                else if (g.size() == 3)
                    for (int q = 0; q < 3; ++q) o->grid[q] = g[q];
                else
                    errors.push_back("--grid takes either 1 or 3 values");
            } else if (key == "--aspect") {
                double b, c;
                if (t.size() != 2 || !parse_dbl(t[0], &b) || !parse_dbl(t[1], &c))
                    errors.push_back("--aspect: expected b/a,c/a, got \"" + v + "\"");
                else {
                    o->aspect_b = b;
                    o->aspect_c = c;
                }
            } else if (key == "--material") {
                double re, im;
                if (t.size() != 2 || !parse_dbl(t[0], &re) || !parse_dbl(t[1], &im))
                    errors.push_back("--material: expected mre,mi, got \"" + v + "\"");
                else {
                    o->mat_re.push_back((float)re);
                    o->mat_im.push_back((float)im);
                }
            } else {
                for (auto& x : t) {
                    if (!parse_int(x, &iv)) {
                        errors.push_back("--seeds: expected integers, got \"" + v + "\"");
                        break;
                    }
                    o->seeds.push_back(iv);
                }
            }
        } else {
            errors.push_back("unknown option \"" + a + "\"");
        }
    }
    if (argc == 1) errors.push_back("no options given (the interactive prompts are not provided)");
    // defaults (options.cpp:625-659)
    if (std::isnan(o->theta_min)) o->theta_min = 0.;
    if (std::isnan(o->theta_max)) o->theta_max = 180.;
    if (std::isnan(o->theta_step)) o->theta_step = 10.;
    if (o->state_file.empty()) o->state_file = "out.dst";
    if (o->output_file.empty()) o->output_file = "out.dat";
    if (o->samples == kUnsetInt) o->samples = 1;
    if (o->orientation == kUnsetInt) o->orientation = 0;
    if (std::isnan(o->psi)) o->psi = 0.;
    if (std::isnan(o->iota)) o->iota = 0.;
    if (std::isnan(o->kappa)) o->kappa = 0.;
    if (o->planes == kUnsetInt) o->planes = 1;
    if (std::isnan(o->aspect_b)) o->aspect_b = 1.;
    if (std::isnan(o->aspect_c)) o->aspect_c = 1.;
    if (std::isnan(o->radius_ratio)) o->radius_ratio = 1.;
    if (o->levels == kUnsetInt) o->levels = 2;
    if (o->bites == kUnsetInt) o->bites = 0;
    if (std::isnan(o->depth)) o->depth = 0.;
    if (o->polarizability == kUnsetInt) o->polarizability = 1;
    if (std::isnan(o->tolerance)) o->tolerance = 1e-3;
    if (o->max_steps == kUnsetInt) o->max_steps = 20;
    if (o->fft_mode == kUnsetInt) o->fft_mode = 0;
    if (o->solver == kUnsetInt) o->solver = 0;
    if (std::isnan(o->tau)) o->tau = 0.;
    if (o->threads == kUnsetInt) o->threads = 0;
    // validation (options.cpp:485-622)
    auto pow2 = [](int n) { return n > 0 && (n & (n - 1)) == 0; };
    const int t = o->target;
    const bool aspect = t == 2 || t == 3 || t == 5 || t == 9 || t == 10 || t == 13;
    const bool seeded = t == 4 || t == 5 || t == 9 || t == 10 || t == 13;
    const bool worm = t == 6 || t == 7 || t == 8;
    const bool single = t == 1 || t == 2 || t == 3 || t == 6 || t == 7 || t == 8 || t == 14;
    const size_t nmat = o->mat_re.size();
    if (!(o->theta_step > 0.)) errors.push_back("--theta-step must be positive");
    if (o->theta_max < o->theta_min) errors.push_back("--theta-max must not be less than --theta-min");
    if (o->samples < 1) errors.push_back("--samples must be at least 1");
    if (o->planes < 1) errors.push_back("--planes must be at least 1");
    if (t < 1 || t > 14) errors.push_back("--target must be 1..14 or a target name");
    {
        bool bad_sign = false, bad_pow2 = false;
        for (int q = 0; q < 3; ++q) {
            if (o->grid[q] <= 0)
                bad_sign = true;
            else if (!pow2(o->grid[q]))
                bad_pow2 = true;
        }
        if (bad_sign)
            errors.push_back("--grid needs 1 or 3 positive dimensions");
        else if (bad_pow2)
            errors.push_back("--grid dimensions must be powers of two");
    }
    if (!(o->size_parameter > 0.)) errors.push_back("--size-parameter must be positive");
    if (!(o->tolerance > 0.)) errors.push_back("--tolerance must be positive");
    if (o->max_steps < 1) errors.push_back("--max-steps must be at least 1");
    if (o->threads < 0) errors.push_back("--threads must be 0 (auto) or positive");
    if (o->group < 1 || o->group > 32) errors.push_back("--group must be between 1 and 32");
    if (o->output_file == o->state_file) errors.push_back("--output and --state must name different files");
    if (o->output_file.size() > 1023) errors.push_back("--output is too long (1024 bytes at most)");
    if (o->state_file.size() > 1023) errors.push_back("--state is too long (1024 bytes at most)");
    if (o->solver == 3 && !(o->tau > 0.)) errors.push_back("solver 3 needs --tau > 0");
    if (t == 11) {
        if (o->shape_file.empty()) errors.push_back("target 11 needs --shape-file");
        if (nmat == 0) errors.push_back("target 11 needs at least one --material");
    } else if (t == 12 || t == 13) {
        if (nmat != 2) errors.push_back("target " + std::to_string(t) + " needs exactly two --material options");
    } else if (single) {
        if (nmat != 1) errors.push_back("this target needs exactly one --material");
    } else if (t >= 4 && t <= 10) {
        if (nmat == 0) errors.push_back("this target needs at least one --material");
    }
    if (seeded && o->seeds.size() != nmat)
        errors.push_back("--seeds needs one value per material (" + std::to_string(nmat) + " expected, " +
                         std::to_string(o->seeds.size()) + " given)");
    if (aspect && (!(o->aspect_b > 0.) || !(o->aspect_c > 0.))) errors.push_back("this target needs --aspect <b/a,c/a>");
    if (t == 12 && !(o->radius_ratio > 0.)) errors.push_back("target 12 needs --radius-ratio > 0");
    if (t == 13 && o->levels < 1) errors.push_back("target 13 needs --levels >= 1");
    if (t == 10 && (o->bites < 0 || !(o->depth >= 0.))) errors.push_back("target 10 needs --bites >= 0 and --depth >= 0");
    if ((worm || t == 14) && !(o->packing >= 0.)) errors.push_back("this target needs --packing >= 0");
    if (worm && !(o->monomer_radius > 0.)) errors.push_back("this target needs --monomer-radius > 0");
    if (worm && !(o->monomer_length_ratio >= 0.)) errors.push_back("this target needs --monomer-length-ratio >= 0");
    if (!errors.empty()) {
        std::fprintf(stderr, "error:\n");
        for (auto& e : errors) std::fprintf(stderr, "  %s\n", e.c_str());
        std::fprintf(stderr, "Try '%s --help' for the list of options.\n", argv[0]);
        return 2;
    }
    if (o->solver != 3 && o->tau > 0.) std::fprintf(stderr, "warning: --tau is only used by solver 3, ignored\n");
    if (o->orientation == 1 && (o->psi != 0. || o->iota != 0. || o->kappa != 0.))
        std::fprintf(stderr, "warning: --psi/--iota/--kappa are only used by --orientation fixed, ignored\n");
    return 0;
}

// rounding() (complex.cpp:213-223)
int round_half_up(float r0) {
    double ip;
    const double fr = std::modf((double)r0, &ip);
    return fr > 0.5 ? (int)ip + 1 : (int)ip;
}

struct State {   // the .dst record (main.cpp:214-262)
    int count = 0;
    float cabs = 0.f, cext = 0.f;
    std::vector<float> mm;   // [nang][16]
};

bool write_state(FILE* f, const State& s) {
    std::fseek(f, 0, SEEK_SET);
    bool ok = std::fwrite(&s.count, sizeof(int), 1, f) == 1;
    ok = ok && std::fwrite(&s.cabs, sizeof(float), 1, f) == 1;
    ok = ok && std::fwrite(&s.cext, sizeof(float), 1, f) == 1;
    ok = ok && std::fwrite(s.mm.data(), sizeof(float), s.mm.size(), f) == s.mm.size();
    std::fflush(f);
    return ok;
}

int fail_exit(int code, const char* what) {
    std::printf("\n %s\n", what);
    std::fflush(stdout);
    return code;
}

}  // namespace

int main(int argc, char** argv) {
    std::printf("\nComputation of light scattering by arbitrary target:");
    std::printf("\nDiscrete Dipole Approximation (DDA) with Fast Fourier Transform (FFT), MI355X\n");
    std::fflush(stdout);
    Options o;
    {
        const int rc = parse(argc, argv, &o);
        if (rc == 1) return 0;
        if (rc != 0) return 2;
    }
    // base seed (randomize.cpp:13-33)
    unsigned long long base = o.seed;
    if (!o.has_seed) {
        base = (unsigned long long)std::time(nullptr);
        if (FILE* f = std::fopen("/dev/random", "rb")) {
            if (std::fread(&base, sizeof(base), 1, f) != 1) base = (unsigned long long)std::time(nullptr);
            std::fclose(f);
        }
    }
    unsigned long long stream = ddafft_rng_worker_seed(base, 0);   // the first worker (-j 1)
    if (stream == 0) stream = 0x9E3779B97F4A7C15ULL;

    const float tmin = (float)o.theta_min, tmax = (float)o.theta_max, tstep = (float)o.theta_step;
    const int nang = round_half_up((float)(((tmax - tmin) / tstep) + 1.));
    const int npl = o.planes, ni = o.grid[0], nj = o.grid[1], nl = o.grid[2];
    const int nd = npl * nang;

    // state file: create or resume (main.cpp:214-262)
    State st;
    st.mm.assign((size_t)nang * 16, 0.f);
    if (o.restart) std::remove(o.state_file.c_str());
    FILE* sf = std::fopen(o.state_file.c_str(), "rb+");
    if (!sf) {
        sf = std::fopen(o.state_file.c_str(), "wb+");
        if (!sf) return fail_exit(2, "Cannot create the dst-file! ");
        if (!write_state(sf, st)) return fail_exit(2, "Cannot create the dst-file! ");
    }
    std::fseek(sf, 0, SEEK_SET);
    if (std::fread(&st.count, sizeof(int), 1, sf) != 1 || std::fread(&st.cabs, sizeof(float), 1, sf) != 1 ||
        std::fread(&st.cext, sizeof(float), 1, sf) != 1 ||
        std::fread(st.mm.data(), sizeof(float), st.mm.size(), sf) != st.mm.size())
        return fail_exit(4, "Cannot read data from the dst-file! ");

    const float x = (float)o.size_parameter;
    const float kd = (float)(2. * x / (float)ni);   // main.cpp:318
    const int nsorts = (int)o.mat_re.size();
    const int nmat = nsorts + 1;
    std::vector<int> mv(nmat, 0);
    for (int q = 0; q < nsorts; ++q) mv[q + 1] = (o.mat_re[q] == 1.f && o.mat_im[q] == 0.f) ? 0 : 1;   // :606-613

    if (st.count < o.samples) {
        // base geometry, rotated once for a fixed orientation (main.cpp:354-387, 516-542)
        std::vector<float> dirs0((size_t)nd * 3), nrm0((size_t)npl * 3), dirsF((size_t)nd * 3), nrmF((size_t)npl * 3);
        float k0[3], kF[3], e0[12], eF[12];
        if (ddafft_geometry_fixed(npl, nang, tmin, tstep, 0., 0., 0., dirs0.data(), nrm0.data(), k0, e0) ||
            ddafft_geometry_fixed(npl, nang, tmin, tstep, o.psi, o.iota, o.kappa, dirsF.data(), nrmF.data(), kF, eF))
            return fail_exit(1, "geometry failed");
        // occupancy
        std::vector<signed char> occ((size_t)ni * nj * nl, 0);
        const int t = o.target;
        const bool fixed_target = !(t == 4 || t == 5 || t == 6 || t == 7 || t == 8 || t == 9 || t == 10 || t == 13);
        const float rw = (float)o.monomer_radius, lw = rw * (float)o.monomer_length_ratio;
        const int nworms = (t >= 6 && t <= 8)
                               ? ddafft_worm_count(ni, nj, nl, o.packing, o.monomer_radius, o.monomer_length_ratio)
                               : 0;
        auto generate = [&]() -> int {
            switch (t) {
                case 1: return ddafft_target_ellipsoid(ni, nj, nl, 1., 1., occ.data());
                case 2: return ddafft_target_ellipsoid(ni, nj, nl, (float)o.aspect_b, (float)o.aspect_c, occ.data());
                case 3: return ddafft_target_parallelepiped(ni, nj, nl, (float)o.aspect_b, (float)o.aspect_c, occ.data());
                case 4:
                    return ddafft_target_random_ellipsoid(ni, nj, nl, 1., 1., nsorts, o.seeds.data(), &stream, occ.data());
                case 5:
                    return ddafft_target_random_ellipsoid(ni, nj, nl, (float)o.aspect_b, (float)o.aspect_c, nsorts,
                                                          o.seeds.data(), &stream, occ.data());
                case 6:
                case 7:
                case 8: return ddafft_target_worms(t, ni, nj, nl, rw, lw, nworms, 0, &stream, occ.data());
                case 9:
                    return ddafft_target_random_parallelepiped(ni, nj, nl, (float)o.aspect_b, (float)o.aspect_c, nsorts,
                                                               o.seeds.data(), &stream, occ.data());
                case 10:
                    return ddafft_target_nibble_ellipsoid(ni, nj, nl, (float)o.aspect_b, (float)o.aspect_c, nsorts,
                                                          o.seeds.data(), o.bites, (float)o.depth, mv.data(), &stream,
                                                          occ.data());
                case 12: return ddafft_target_bisphere(ni, nj, nl, (float)o.radius_ratio, occ.data());
                case 13:
                    return ddafft_target_hierarchical_ellipsoid(o.levels, ni, nj, nl, (float)o.aspect_b,
                                                                (float)o.aspect_c, nsorts, o.seeds.data(), &stream,
                                                                occ.data());
                case 14: return ddafft_target_sphere_spec(ni, nj, nl, ni / 2., (float)o.packing, occ.data());
            }
            return 0;
        };
        if (t == 11) {
            const int rc = ddafft_target_file_read(o.shape_file.c_str(), ni, nj, nl, occ.data());
            if (rc == -3)
                return fail_exit(3, "Sample particle in the file is out of the expected boundaries.");
            if (rc < 0) return fail_exit(2, "Cannot read the initial data file!");
            for (size_t k = 0; k < occ.size(); ++k)
                if (occ[k] < 0 || occ[k] >= nmat) return fail_exit(3, "material index in the file outside the material list");
        } else if (fixed_target) {
            if (generate() < 0) return fail_exit(1, "the target does not fit the lattice");
        }
        // polarizability per material and polarization (main.cpp:602-631, 926-931)
        auto alpha_of = [&](const float* kinc, const float* einc, std::vector<float>& out) {
            out.assign((size_t)nmat * 2, 0.f);
            for (int q = 0; q < nsorts; ++q) {
                const float m[2] = {o.mat_re[q], o.mat_im[q]};
                ddafft_polarizability(o.polarizability, kd, m, kinc, einc, &out[2 * (q + 1)]);
            }
        };
        // Samples go to the GPU in groups: a fixed target's Monte-Carlo
        // orientations as right-hand sides of one solve; random targets (a new
        // realization each sample) as instances of one batched solve, one
        // context each (their number capped so that they fit HBM comfortably).
        const long long ncell = (long long)ni * nj * nl;
        const int group = fixed_target ? (o.orientation == 1 ? o.group : 1)
                                       : std::max(1, (int)std::min<long long>(o.group, (1LL << 24) / ncell));
        const int nctx = fixed_target ? 1 : group;
        std::vector<ddafft_ctx*> ctxs(nctx, nullptr);
        for (int q = 0; q < nctx; ++q)
            if (ddafft_create(ni, nj, nl, kd, fixed_target ? 2 * group : 2, &ctxs[q]) != DDAFFT_OK ||
                (o.reference_sums && ddafft_set_sum_mode(ctxs[q], 1) != DDAFFT_OK) ||
                (o.compact_ahat && ddafft_set_ahat_mode(ctxs[q], 1) != DDAFFT_OK)) {
                std::fprintf(stderr, "%s\n", ddafft_last_error());
                return 1;
            }
        ddafft_ctx* ctx = ctxs[0];
        bool target_set = false;
        std::vector<float> esc;
        // fixed target and fixed orientation: every sample is the same solve
        std::vector<float> cached_mm;
        float cached_sum[2] = {0.f, 0.f};
        bool have_cache = false;
        while (st.count < o.samples) {
            const int g = std::min(group, o.samples - st.count);
            if (fixed_target && !target_set) {
                if (ddafft_set_target(ctx, occ.data(), nmat, mv.data()) != DDAFFT_OK) {
                    std::fprintf(stderr, "%s\n", ddafft_last_error());
                    return 1;
                }
                target_set = true;
            }
            std::vector<float> sum(2 * (size_t)g), mm((size_t)g * nang * 16);
            if (o.orientation == 0 && fixed_target && have_cache) {
                for (int s = 0; s < g; ++s) {
                    sum[2 * s] = cached_sum[0];
                    sum[2 * s + 1] = cached_sum[1];
                    std::memcpy(&mm[(size_t)s * nang * 16], cached_mm.data(), cached_mm.size() * sizeof(float));
                }
            } else {
                // incidence of each sample: fixed, or a Monte-Carlo draw (main.cpp:875-921)
                const int nrhs = 2 * g;
                std::vector<float> kin((size_t)nrhs * 3), ein((size_t)nrhs * 6, 0.f), dr((size_t)nrhs * nd * 3),
                    alpha((size_t)nrhs * nmat * 2), nrm((size_t)g * npl * 3), kk((size_t)g * 3), ee((size_t)g * 12);
                for (int s = 0; s < g; ++s) {
                    float* ks = &kk[3 * s];
                    float* es = &ee[12 * s];
                    float* ns = &nrm[(size_t)s * npl * 3];
                    std::vector<float> ds((size_t)nd * 3);
                    // a random target's realization comes first in the sample (main.cpp:813-857)
                    if (!fixed_target) {
                        if (generate() < 0) return fail_exit(1, "the target does not fit the lattice");
                        if (ddafft_set_target(ctxs[s], occ.data(), nmat, mv.data()) != DDAFFT_OK) {
                            std::fprintf(stderr, "%s\n", ddafft_last_error());
                            return 1;
                        }
                    }
                    if (o.orientation == 1) {
                        float rot[9];
                        ddafft_orientation_mc(&stream, rot);
                        ddafft_rotate(rot, nd, dirs0.data(), ds.data());
                        ddafft_rotate(rot, npl, nrm0.data(), ns);
                        ddafft_rotate(rot, 1, k0, ks);
                        const float ex[3] = {e0[0], e0[2], e0[4]}, ey[3] = {e0[6], e0[8], e0[10]};
                        float rx[3], ry[3];
                        ddafft_rotate(rot, 1, ex, rx);
                        ddafft_rotate(rot, 1, ey, ry);
                        for (int q = 0; q < 3; ++q) {
                            es[2 * q] = rx[q];
                            es[6 + 2 * q] = ry[q];
                        }
                    } else {
                        ds = dirsF;
                        std::memcpy(ns, nrmF.data(), npl * 3 * sizeof(float));
                        std::memcpy(ks, kF, sizeof(kF));
                        std::memcpy(es, eF, sizeof(eF));
                    }
                    for (int p = 0; p < 2; ++p) {
                        const int r = 2 * s + p;
                        std::memcpy(&kin[3 * r], ks, 3 * sizeof(float));
                        std::memcpy(&ein[6 * r], es + 6 * p, 6 * sizeof(float));
                        std::memcpy(&dr[(size_t)r * nd * 3], ds.data(), ds.size() * sizeof(float));
                        std::vector<float> a;
                        alpha_of(ks, es + 6 * p, a);
                        std::memcpy(&alpha[(size_t)r * nmat * 2], a.data(), a.size() * sizeof(float));
                    }
                }
                esc.assign((size_t)nrhs * nd * 6, 0.f);
                std::vector<float> cabs(nrhs), cext(nrhs);
                std::vector<int> info((size_t)nrhs * 4);
                ddafft_job j{};
                j.nrhs = nrhs;
                j.alpha = alpha.data();
                j.kinc = kin.data();
                j.einc = ein.data();
                j.ndirs = nd;
                j.dirs = dr.data();
                j.dirs_per_rhs = 1;
                j.esc = esc.data();
                j.cabs = cabs.data();
                j.cext = cext.data();
                j.info = info.data();
                j.exact_phases = 1;   // the reference's glibc phases for Monte-Carlo incidences too
                int rc = DDAFFT_OK;
                if (fixed_target) {
                    rc = ddafft_solve_ex(ctx, &j, o.solver, (float)o.tau, o.max_steps, (float)o.tolerance, nullptr);
                } else {
                    // one instance per sample: its two polarizations
                    std::vector<ddafft_job> jobs(g, j);
                    for (int s = 0; s < g; ++s) {
                        ddafft_job& js = jobs[s];
                        js.nrhs = 2;
                        js.alpha = &alpha[(size_t)2 * s * nmat * 2];
                        js.kinc = &kin[(size_t)2 * s * 3];
                        js.einc = &ein[(size_t)2 * s * 6];
                        js.dirs = &dr[(size_t)2 * s * nd * 3];
                        js.esc = &esc[(size_t)2 * s * nd * 6];
                        js.cabs = &cabs[2 * s];
                        js.cext = &cext[2 * s];
                        js.info = &info[(size_t)2 * s * 4];
                    }
                    rc = ddafft_solve_batch(g, ctxs.data(), jobs.data(), o.solver, (float)o.tau, o.max_steps,
                                            (float)o.tolerance, g);
                }
                if (rc != DDAFFT_OK) {
                    std::fprintf(stderr, "%s\n", ddafft_last_error());
                    return 1;
                }
                for (int s = 0; s < g; ++s) {
                    // rCabsSum += .5*rCabs per polarization (main.cpp:939)
                    float ca = 0.f, ce = 0.f;
                    for (int p = 0; p < 2; ++p) {
                        ca = (float)(ca + .5 * cabs[2 * s + p]);
                        ce = (float)(ce + .5 * cext[2 * s + p]);
                    }
                    sum[2 * s] = ca;
                    sum[2 * s + 1] = ce;
                    ddafft_muller_average(npl, nang, &kk[3 * s], &dr[(size_t)(2 * s) * nd * 3], &nrm[(size_t)s * npl * 3],
                                          &ee[12 * s], &esc[(size_t)(2 * s) * nd * 6],
                                          &esc[(size_t)(2 * s + 1) * nd * 6], &mm[(size_t)s * nang * 16]);
                    std::printf("\n Sample #%d: steps %d/%d, Cabs %.6f, Cext %.6f", st.count + s + 1, info[8 * s],
                                info[8 * s + 4], ca, ce);
                }
                if (o.orientation == 0 && fixed_target) {
                    cached_sum[0] = sum[0];
                    cached_sum[1] = sum[1];
                    cached_mm.assign(mm.begin(), mm.begin() + (size_t)nang * 16);
                    have_cache = true;
                }
            }
            // accumulate in sample order and refresh the state file (main.cpp:958-983)
            for (int s = 0; s < g; ++s) {
                st.cabs = st.cabs + sum[2 * s];
                st.cext = st.cext + sum[2 * s + 1];
                for (size_t q = 0; q < (size_t)nang * 16; ++q) st.mm[q] = st.mm[q] + mm[(size_t)s * nang * 16 + q];
                st.count++;
                if (!write_state(sf, st)) return fail_exit(2, "Cannot write the dst-file! ");
            }
            std::fflush(stdout);
        }
        for (ddafft_ctx* q : ctxs) ddafft_destroy(q);
    }
    std::fclose(sf);

    // out.dat (main.cpp:1085-1118)
    std::printf("\n Averaging over %d samples is done.", st.count);
    FILE* out = std::fopen(o.output_file.c_str(), "w");
    if (!out) return fail_exit(2, "Cannot read the output data file!");
    const float w = (float)(1. / (float)st.count);
    std::fprintf(out, "Cabs=%.8f\t\tCext=%.8f", (float)(w * st.cabs), (float)(w * st.cext));
    for (int m = 0; m < nang; ++m) {
        const float theta = tmin + tstep * (float)m;
        std::fprintf(out, "\n%7.3f\t", (float)theta);
        for (int q = 0; q < 16; ++q) std::fprintf(out, "\t%+10.8e", (float)(w * st.mm[(size_t)m * 16 + q]));
    }
    std::fclose(out);
    std::printf("\n");
    return 0;
}
