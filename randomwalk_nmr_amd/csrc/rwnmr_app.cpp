// rwnmr_app.cpp -- the RWNMR command-line program on the MI355X library:
// reads the reference's configuration files, loads the micro-CT image, sets
// up the walkers and runs the CPMG, PFGSE and multitau sequences on the GPU,
// writing the reference's output tree under <root>/db/<NAME>/.
//
//   rwnmr [-rwconfig FILE] [-uctconfig FILE] cpmg [-config FILE]
//         pfgse [-config FILE] multitau [-config MT CPMG | -mtconfig MT | -cpmgconfig CPMG] ga
//
// Reference behaviour restated (paths under src/backend/ of RandomWalk_NMR):
//   argument grammar                     B/Utils/ArgsParser.cpp:10-151
//   essentials, command loop             B/App/rwnmrApp.cpp:26-150
//   configuration files (defaults first) B/ConfigFiles/*_config.cpp (constructors)
//   simulation set-up and info files     B/NMR_Simulation/NMR_Simulation.cpp:44-483, 1588-1712
//   CPMG sequence and its files          B/NMR_Simulation/NMR_cpmg.cpp:23-636
//   PFGSE sequence and its files         B/NMR_Simulation/NMR_pfgse.cpp:30-1656
//   multitau sequence and its files      B/NMR_Simulation/NMR_multitau.cpp:23-346
//
// The project root is $RWNMR_ROOT (default: the working directory): the
// configuration root is <root>./config/ and results go to <root>./db/.  Keys
// a configuration file leaves unset take the values of the reference's
// config/default files (also where the reference leaves the member
// uninitialised); a config/default file under the configuration root, when
// present, is read first as the reference does.  The walk always runs on the
// GPU (GPU_USAGE is reported, not obeyed): this program has no CPU path.
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>

#include <algorithm>
#include <map>
#include <memory>
#include <random>
#include <string>
#include <vector>

#include "../../include/rwnmr/rwnmr.h"
#include "../../include/rwnmr/rwnmr_config.h"
#include "../../include/rwnmr/rwnmr_drivers.h"
#include "../../include/rwnmr/rwnmr_host.h"
#include "../../include/rwnmr/rwnmr_io.h"
#include "../../include/rwnmr/rwnmr_output.h"

namespace {

[[noreturn]] void die(const std::string &msg)
{
    fprintf(stderr, "rwnmr: %s\n", msg.c_str());
    exit(1);
}

void check(rwnmr_status s, const std::string &what, const rwnmr_ctx *ctx = nullptr)
{
    if (s != RWNMR_OK)
        die(what + " failed (status " + std::to_string(s) + ")" +
            (ctx ? std::string(": ") + rwnmr_last_error(ctx) : std::string()));
}

bool file_exists(const std::string &p)
{
    struct stat st;
    return stat(p.c_str(), &st) == 0 && S_ISREG(st.st_mode);
}

// ---------------------------------------------------------------- arguments

// ArgsParser (B/Utils/ArgsParser.cpp:10-151): commands[0..1] are always
// "rwnmr" and "uct"; every argument is inspected in turn, so a path equal to
// a command name also starts that command, as in the reference.
struct Args {
    std::vector<std::string> commands{"rwnmr", "uct"};
    std::vector<std::string> paths{"default", "default"};

    Args(int argc, char **argv)
    {
        for (int i = 1; i < argc; ++i) {
            const std::string a = argv[i];
            auto next = [&](int k) { return i + k < argc ? std::string(argv[i + k]) : std::string(); };
            if (a == "-rwconfig") {
                if (i + 1 < argc)
                    paths[0] = next(1);
            } else if (a == "-uctconfig") {
                if (i + 1 < argc)
                    paths[1] = next(1);
            } else if (a == "cpmg" || a == "pfgse" || a == "ga") {
                commands.push_back(a);
                paths.push_back(i + 1 < argc && next(1) == "-config" && i + 2 < argc ? next(2) : "default");
            } else if (a == "multitau") {
                commands.push_back("multitau");
                commands.push_back("mtoff");
                std::string mt = "default", cp = "default";
                if (i + 1 < argc) {
                    const std::string f = next(1);
                    if (f == "-config" && i + 3 < argc) {
                        mt = next(2);
                        cp = next(3);
                    } else if (f == "-mtconfig" && i + 2 < argc) {
                        mt = next(2);
                    } else if (f == "-cpmgconfig" && i + 2 < argc) {
                        cp = next(2);
                    }
                }
                paths.push_back(mt);
                paths.push_back(cp);
            }
        }
    }
};

// ---------------------------------------------------------------- configuration

// A configuration: the reference's default values as fallbacks, then the
// files read in order (rwnmr_config_read keeps every value; the last wins,
// lists of RHO append).
struct Conf {
    rwnmr_config_file *c = nullptr;
    std::map<std::string, std::string> fallback;

    Conf(std::initializer_list<std::pair<const std::string, std::string>> defaults) : fallback(defaults)
    {
        check(rwnmr_config_create(&c), "config");
    }
    ~Conf() { rwnmr_config_destroy(c); }
    Conf(const Conf &) = delete;
    Conf &operator=(const Conf &) = delete;

    void read(const std::string &path, bool required)
    {
        if (!file_exists(path)) {
            if (required)
                die("could not open config file " + path);
            return;
        }
        check(rwnmr_config_read(c, path.c_str()), "reading " + path);
    }
    bool has(const char *k) const { return rwnmr_config_has(c, k) || fallback.count(k); }
    std::string s(const char *k) const
    {
        if (rwnmr_config_has(c, k)) {
            std::vector<char> buf(4096);
            check(rwnmr_config_string(c, k, buf.data(), (uint32_t)buf.size()), std::string("key ") + k);
            return buf.data();
        }
        auto it = fallback.find(k);
        return it == fallback.end() ? std::string() : it->second;
    }
    double d(const char *k) const
    {
        const std::string v = s(k);
        try {
            return std::stod(v);
        } catch (...) {
            die(std::string("key ") + k + ": not a number: '" + v + "'");
        }
    }
    int64_t i(const char *k) const
    {
        const std::string v = s(k);
        try {
            return std::stoll(v);
        } catch (...) {
            die(std::string("key ") + k + ": not an integer: '" + v + "'");
        }
    }
    bool b(const char *k) const { return s(k) == "true"; }
    // "{a, b, c}" or one number; append: every value read (RHO)
    std::vector<double> list(const char *k, bool append) const
    {
        uint32_t n = 0;
        if (rwnmr_config_has(c, k)) {
            check(rwnmr_config_list(c, k, append ? 1 : 0, nullptr, 0, &n), std::string("list ") + k);
            std::vector<double> out(n);
            check(rwnmr_config_list(c, k, append ? 1 : 0, out.data(), n, &n), std::string("list ") + k);
            return out;
        }
        rwnmr_config_file *tmp = nullptr;
        check(rwnmr_config_create(&tmp), "config");
        auto it = fallback.find(k);
        std::vector<double> out;
        if (it != fallback.end()) {
            rwnmr_config_set(tmp, k, it->second.c_str());
            if (rwnmr_config_list(tmp, k, 0, nullptr, 0, &n) == RWNMR_OK) {
                out.resize(n);
                rwnmr_config_list(tmp, k, 0, out.data(), n, &n);
            }
        }
        rwnmr_config_destroy(tmp);
        return out;
    }
};

// config/default/rwnmr.config (plus GIROMAGNETIC_RATIO / _UNIT, SAVE_WALKERS
// and HISTOGRAM_SCALE, which that file omits)
Conf *rwnmr_conf()
{
    return new Conf({{"NAME", "NMR_simulation"}, {"D0", "2.5"}, {"BULK_TIME", "2800"}, {"RHO_TYPE", "uniform"},
                     {"RHO", "{0.0}"}, {"WALKERS", "50000"}, {"WALKER_SAMPLES", "1"},
                     {"WALKERS_PLACEMENT", "random"}, {"PLACEMENT_DEVIATION", "0"}, {"STEPS_PER_ECHO", "1"},
                     {"SEED", "0"}, {"BC", "noflux"}, {"SAVE_IMG_INFO", "true"}, {"SAVE_BINIMG", "false"},
                     {"SAVE_WALKERS", "false"}, {"HISTOGRAMS", "1"}, {"HISTOGRAM_SIZE", "1024"},
                     {"HISTOGRAM_SCALE", "linear"}, {"GPU_USAGE", "true"}, {"GIROMAGNETIC_RATIO", "42.576"},
                     {"GIROMAGNETIC_UNIT", "mhertz"}});
}

// config/default/cpmg.config
Conf *cpmg_conf()
{
    return new Conf({{"METHOD", "image-based"}, {"OBS_TIME", "3000.0"}, {"APPLY_BULK", "true"},
                     {"RESIDUAL_FIELD", "none"}, {"GRADIENT_VALUE", "0.0"}, {"GRADIENT_DIRECTION", "2"},
                     {"MIN_T2", "0.1"}, {"MAX_T2", "10000.0"}, {"USE_T2_LOGSPACE", "true"}, {"NUM_T2_BINS", "128"},
                     {"MIN_LAMBDA", "-4"}, {"MAX_LAMBDA", "2"}, {"NUM_LAMBDAS", "512"}, {"PRUNE_NUM", "512"},
                     {"NOISE_AMP", "0.0"}, {"SAVE_MODE", "true"}, {"SAVE_T2", "true"}, {"SAVE_WALKERS", "false"},
                     {"SAVE_DECAY", "true"}, {"SAVE_HISTOGRAM", "true"}, {"SAVE_HISTOGRAM_LIST", "false"}});
}

// config/default/pfgse.config (+ THRESHOLD_WINDOW 1, NOISE_AMP 0,
// TARGET_SNR -1 of the pfgse_config constructor)
Conf *pfgse_conf()
{
    return new Conf({{"APPLY_BULK", "true"}, {"PULSE_WIDTH", "0.1"}, {"MAX_GRADIENT", "{3000.0, 0.0, 0.0}"},
                     {"GRADIENT_SAMPLES", "10"}, {"TIME_SEQ", "log"}, {"TIME_SAMPLES", "3"}, {"TIME_MIN", "-2.0"},
                     {"TIME_MAX", "0.3"}, {"APPLY_SCALE_FACTOR", "true"}, {"INSPECTION_LENGTH", "100.0"},
                     {"THRESHOLD_TYPE", "samples"}, {"THRESHOLD_VALUE", "8"}, {"THRESHOLD_WINDOW", "1"},
                     {"NOISE_AMP", "0.0"}, {"TARGET_SNR", "-1.0"}, {"ALLOW_WALKER_SAMPLING", "true"},
                     {"APPLY_ABSORPTION", "false"}, {"SAVE_MODE", "true"}, {"SAVE_PFGSE", "true"},
                     {"SAVE_WALKERS", "false"}, {"SAVE_HISTOGRAM", "true"}, {"SAVE_HISTOGRAM_LIST", "false"}});
}

// config/default/multitau.config (+ COMPLETE_DECAY false)
Conf *multitau_conf()
{
    return new Conf({{"TAU_MIN", "0.01"}, {"TAU_MAX", "10.0"}, {"TAU_POINTS", "10"}, {"TAU_SCALE", "log"},
                     {"COMPLETE_DECAY", "false"}, {"SAVE_MODE", "true"}, {"SAVE_DECAY", "true"},
                     {"SAVE_T2", "true"}, {"SAVE_WALKERS", "true"}, {"SAVE_HISTOGRAM", "true"},
                     {"SAVE_HISTOGRAM_LIST", "false"}});
}

// std::to_string(v) cut `decimals` places after the point of std::to_string(ref)
// (NMR_cpmg::setName, NMR_multitau::setName / setCPMG)
std::string trimmed(double v, int decimals, double ref)
{
    const std::string s = std::to_string(v);
    return s.substr(0, std::to_string(ref).find('.') + decimals + 1);
}

// ---------------------------------------------------------------- simulation state

struct Walkers {
    uint64_t n = 0;
    std::vector<int32_t> x0, y0, z0, x, y, z;
    std::vector<uint64_t> seed0, seed;
    std::vector<double> energy, xi, penalty;
    std::vector<uint32_t> collisions;

    void reset_walk()     // resetPosition + resetSeed (walker.h)
    {
        x = x0;
        y = y0;
        z = z0;
        seed = seed0;
    }
    void reset_energy() { std::fill(energy.begin(), energy.end(), 1.0); }
    rwnmr_walkers view(bool with_hits)
    {
        rwnmr_walkers w;
        memset(&w, 0, sizeof(w));
        w.x = x.data();
        w.y = y.data();
        w.z = z.data();
        w.seed = seed.data();
        w.energy = energy.data();
        w.hits = with_hits ? collisions.data() : nullptr;
        w.penalty = penalty.data();
        w.n = n;
        w.on_device = 0;
        return w;
    }
};

struct Histograms {   // the collision histograms of the last map walk
    int32_t nb = 0, n_used = 0;
    std::vector<double> bins, amps;            // global
    std::vector<double> list_bins, list_amps;  // [n_used][nb]
};

struct Sim {
    std::string root, config_root, db_path, name, sim_dir;
    // rwnmr.config
    double D0 = 0, bulk_time = 0, gamma = 0;
    std::string rho_type, bc_name, placement;
    std::vector<double> rho;
    uint64_t n_walkers = 0, seed = 0;
    uint32_t samples = 1, steps_per_echo = 1;
    int32_t bc = 0, deviation = 0, histograms = 1, histogram_size = 1024;
    bool hist_log = false, gpu = true;
    // image
    rwnmr_uct_config uct;
    int32_t W = 0, H = 0, D = 0;
    std::vector<uint8_t> solid;
    std::vector<uint64_t> blocks;
    uint64_t pores = 0, interfaces = 0;
    double porosity = 0, svp = 0, resolution = 0, voxel = 0, occupancy = 0;
    uint32_t shift = 0;
    std::string image_path;
    // walkers, generator, time framework
    Walkers w;
    std::mt19937 rng;
    uint64_t sim_steps = 0;
    uint32_t echoes = 0;
    Histograms hist;
    rwnmr_ctx *ctx = nullptr;
    rwnmr_image *img = nullptr;

    double dt() const { return rwnmr_time_step(voxel, D0); }

    // setTimeFramework(double) (NMR_Simulation.cpp:276-298)
    void time_framework(double time_ms)
    {
        check(rwnmr_time_framework(time_ms, voxel, D0, steps_per_echo, &echoes, &sim_steps), "time framework");
    }

    // walker relaxivity from its collision rate (updateWalkersRelaxativity,
    // NMR_Simulation.cpp:1500-1518; B/Walker/walker.h:170-209)
    void update_relaxivity()
    {
        for (uint64_t i = 0; i < w.n; ++i) {
            w.xi[i] = (double)w.collisions[i] / (double)sim_steps;
            double r = rho.empty() ? 0.0 : rho[0];
            if (rho_type == "sigmoid") {
                if (rho.size() < 8)
                    die("RHO_TYPE sigmoid needs 8 RHO values");
                const double *p = rho.data();
                const double xi = w.xi[i];
                r = (p[1] + (p[0] - p[1]) / (1 + exp((-1) * p[3] * (xi - p[2])))) +
                    (p[5] + (p[4] - p[5]) / (1 + exp((-1) * p[7] * (xi - p[6]))));
            }
            w.penalty[i] = rwnmr_penalty(r, voxel, D0);
        }
    }

    // mapSimulation_CUDA_3D_histograms (NMR_Simulation_cuda.cpp:726-1270):
    // walkers restart from their initial state with no collisions and walk
    // sim_steps steps in the segments of the histogram list; a histogram per
    // segment, one of the total collisions over all steps.
    void map_walk()
    {
        w.reset_walk();
        std::fill(w.collisions.begin(), w.collisions.end(), 0u);
        std::vector<uint32_t> first((size_t)histograms), last((size_t)histograms);
        int32_t nh = 0;
        check(rwnmr_histogram_list(echoes, histograms, &nh, first.data(), last.data()), "histogram list");
        const int32_t nb = rwnmr_histogram_bins(histogram_size, hist_log ? 1 : 0);
        hist.nb = nb;
        hist.n_used = nh;
        hist.list_bins.assign((size_t)nh * nb, 0.0);
        hist.list_amps.assign((size_t)nh * nb, 0.0);
        std::vector<uint32_t> seg(w.n);
        for (int32_t h = 0; h < nh; ++h) {
            const uint64_t steps = (uint64_t)steps_per_echo * (last[h] - first[h]);
            std::fill(seg.begin(), seg.end(), 0u);
            rwnmr_walkers v = w.view(false);
            v.energy = nullptr;
            v.penalty = nullptr;
            v.hits = seg.data();
            check(rwnmr_map_run(ctx, img, &v, steps, bc, shift), "map walk", ctx);
            check(rwnmr_collision_histogram(seg.data(), w.n, steps, histogram_size, hist_log ? 1 : 0,
                                            hist.list_bins.data() + (size_t)h * nb,
                                            hist.list_amps.data() + (size_t)h * nb),
                  "segment histogram");
            for (uint64_t i = 0; i < w.n; ++i)
                w.collisions[i] += seg[i];
        }
        hist.bins.assign(nb, 0.0);
        hist.amps.assign(nb, 0.0);
        check(rwnmr_collision_histogram(w.collisions.data(), w.n, sim_steps, histogram_size, hist_log ? 1 : 0,
                                        hist.bins.data(), hist.amps.data()),
              "histogram");
    }

    void write_walkers(const std::string &path)
    {
        check(rwnmr_write_walkers_csv(path.c_str(), w.n, w.x0.data(), w.y0.data(), w.z0.data(), w.x.data(),
                                      w.y.data(), w.z.data(), w.collisions.data(), w.xi.data(), w.energy.data(),
                                      w.seed0.data()),
              "writing " + path);
    }
    void write_histogram(const std::string &path, const std::vector<double> &bins, const std::vector<double> &amps)
    {
        check(rwnmr_write_two_columns(path.c_str(), "Bins,Amps", (uint32_t)bins.size(), bins.data(), amps.data()),
              "writing " + path);
    }
    void write_histogram_list(const std::string &path)
    {
        check(rwnmr_write_histogram_list(path.c_str(), hist.n_used, hist.nb, hist.list_bins.data(),
                                         hist.list_amps.data()),
              "writing " + path);
    }
};

// ---------------------------------------------------------------- essentials

// rwnmrApp::buildEssentials (B/App/rwnmrApp.cpp:38-72) with the
// NMR_Simulation constructor (NMR_Simulation.cpp:44-108), readImage
// (:300-322), setWalkers (:324-363) and save (:412-447).
void build_essentials(Sim &S, const Args &args)
{
    std::unique_ptr<Conf> rc(rwnmr_conf());
    const std::string rw_default = S.config_root + "./default/rwnmr.config";
    const std::string rw_user = args.paths[0] != "default" ? S.config_root + args.paths[0] : rw_default;
    rc->read(rw_default, false);
    if (rw_user != rw_default)
        rc->read(rw_user, true);

    S.name = rc->s("NAME");
    S.D0 = rc->d("D0");
    S.bulk_time = rc->d("BULK_TIME");
    S.rho_type = rc->s("RHO_TYPE");
    S.rho = rc->list("RHO", true);
    S.n_walkers = (uint64_t)rc->i("WALKERS");
    S.samples = (uint32_t)rc->i("WALKER_SAMPLES");
    S.placement = rc->s("WALKERS_PLACEMENT");
    S.deviation = (int32_t)rc->i("PLACEMENT_DEVIATION");
    S.steps_per_echo = (uint32_t)rc->i("STEPS_PER_ECHO");
    S.bc_name = rc->s("BC");
    S.histograms = (int32_t)rc->i("HISTOGRAMS");
    S.histogram_size = (int32_t)rc->i("HISTOGRAM_SIZE");
    S.hist_log = rc->s("HISTOGRAM_SCALE") == "log";
    S.gpu = rc->b("GPU_USAGE");
    // setGiromagneticRatio (NMR_Simulation.cpp:158-162): MHz/T times 2 pi
    S.gamma = rc->d("GIROMAGNETIC_RATIO");
    if (rc->s("GIROMAGNETIC_UNIT") == "mhertz")
        S.gamma *= 6.2831853071795864769252867665590057683943;
    // rwnmr_config::readBC: "periodic" / "mirror", anything else noflux
    S.bc = S.bc_name == "periodic" ? RWNMR_BC_PERIODIC : S.bc_name == "mirror" ? RWNMR_BC_MIRROR : RWNMR_BC_NOFLUX;
    if (S.bc == RWNMR_BC_NOFLUX)
        S.bc_name = "noflux";
    if (S.steps_per_echo == 0 || S.n_walkers == 0 || S.samples == 0)
        die("WALKERS, WALKER_SAMPLES and STEPS_PER_ECHO must be positive");
    S.seed = (uint64_t)rc->i("SEED");
    if (S.seed == 0) {
        // myRNG::RNG_uint64: a random 64-bit seed
        std::random_device rd;
        S.seed = ((uint64_t)rd() << 32) | (uint64_t)rd();
    }
    if (!S.gpu)
        printf("GPU_USAGE false: this program runs the walk on the GPU only\n");

    // uct.config: defaults, config/default/uct.config, the user's file
    rwnmr_uct_config_defaults(&S.uct);
    const std::string uct_default = S.config_root + "./default/uct.config";
    const std::string uct_user = args.paths[1] != "default" ? S.config_root + args.paths[1] : uct_default;
    if (file_exists(uct_default))
        check(rwnmr_uct_config_read(uct_default.c_str(), &S.uct), "reading " + uct_default);
    if (uct_user != uct_default) {
        if (!file_exists(uct_user))
            die("could not open config file " + uct_user);
        check(rwnmr_uct_config_read(uct_user.c_str(), &S.uct), "reading " + uct_user);
    }
    S.resolution = S.uct.resolution;
    S.shift = S.uct.voxel_division > 0 ? (uint32_t)S.uct.voxel_division : 0u;
    S.voxel = S.resolution;

    // results directory <root>./db/<NAME>
    S.db_path = S.root + "./db/";
    S.sim_dir = S.db_path + S.name;
    check(rwnmr_make_dirs(S.sim_dir.c_str()), "creating " + S.sim_dir);

    printf("-- Loading uCT-image\n");
    check(rwnmr_uct_dims(&S.uct, &S.W, &S.H, &S.D), "image size");
    S.solid.resize((size_t)S.W * S.H * S.D);
    check(rwnmr_uct_load(&S.uct, S.W, S.H, S.D, S.solid.data()), "image load");
    const int32_t bC = (S.W + 3) / 4, bR = (S.H + 3) / 4, bD = (S.D + 3) / 4;
    S.blocks.resize((size_t)bC * bR * bD);
    check(rwnmr_bitblock_pack(S.solid.data(), S.W, S.H, S.D, S.blocks.data()), "bitblock pack");
    check(rwnmr_image_stats(S.solid.data(), S.W, S.H, S.D, S.voxel, &S.pores, &S.porosity, &S.interfaces, &S.svp),
          "image statistics");
    // ImagePath::updateCompletePath (B/Utils/ImagePath.cpp): first slice
    // named with as many digits as FIRST_IDX + SLICES has
    {
        int final_id = S.uct.first_idx + S.uct.slices, digits = 1;
        for (int r = final_id / 10; r > 0; r /= 10)
            ++digits;
        char id[64];
        snprintf(id, sizeof(id), "%0*d", digits, S.uct.first_idx);
        S.image_path = std::string(S.uct.dir_path) + S.uct.filename + id + S.uct.extension;
    }
    printf("%llu pore voxel(s), porosity %g, S/Vp %g\n", (unsigned long long)S.pores, S.porosity, S.svp);

    check(rwnmr_ctx_create(&S.ctx, 0), "GPU context");
    check(rwnmr_image_upload(S.ctx, S.blocks.data(), S.W, S.H, S.D, nullptr, 0, &S.img), "image upload", S.ctx);

    printf("-- Setting random walkers\n");
    Walkers &w = S.w;
    w.n = S.n_walkers;
    for (auto *v : {&w.x0, &w.y0, &w.z0, &w.x, &w.y, &w.z})
        v->assign(w.n, 0);
    w.seed0.assign(w.n, 0);
    w.energy.assign(w.n, 1.0);
    w.xi.assign(w.n, 0.0);
    w.penalty.assign(w.n, 1.0);
    w.collisions.assign(w.n, 0u);
    const int32_t mode = S.placement == "point" ? 1 : S.placement == "cubic" ? 2 : 0;
    uint64_t draws = 0;
    check(rwnmr_setup_walkers(S.solid.data(), S.W, S.H, S.D, S.seed, w.n, mode, S.deviation, S.shift,
                              w.seed0.data(), w.x0.data(), w.y0.data(), w.z0.data(), &S.occupancy, &draws),
          "walker placement (" + S.placement + ")");
    w.reset_walk();
    // the simulation generator continues after the placement draws
    S.rng.seed((uint32_t)S.seed);
    S.rng.discard(draws);
    if (S.shift > 0)
        S.voxel = S.resolution / (double)(1u << S.shift);   // applyVoxelDivision :192-204

    // save (NMR_Simulation.cpp:412-447)
    check(rwnmr_write_sim_info((S.sim_dir + "/SimInfo.txt").c_str(), S.name.c_str(), S.sim_dir.c_str(),
                               S.image_path.c_str(), S.voxel, S.D0, S.uct.slices, S.occupancy, S.seed, S.gpu ? 1 : 0,
                               S.bc_name.c_str()),
          "SimInfo.txt");
    if (rc->b("SAVE_IMG_INFO"))
        check(rwnmr_write_image_info((S.sim_dir + "/ImageInfo.txt").c_str(), S.image_path.c_str(), S.W, S.H, S.D,
                                     S.resolution, (int32_t)(1u << S.shift), S.voxel, S.porosity, S.svp),
              "ImageInfo.txt");
    if (rc->b("SAVE_BINIMG"))
        check(rwnmr_write_bin_image((S.sim_dir + "/NMR_binImage.txt").c_str(), S.blocks.data(), S.W, S.H, S.D),
              "NMR_binImage.txt");
    if (rc->b("SAVE_WALKERS"))
        S.write_walkers(S.sim_dir + "/walkers.csv");
}

// ---------------------------------------------------------------- CPMG

struct Cpmg {
    Sim &S;
    Conf &C;
    double exposure;
    std::string dir;
    std::vector<double> times, amps, noise, t2_bins, t2_amps;
    std::vector<double> field;

    Cpmg(Sim &s, Conf &c) : S(s), C(c), exposure(c.d("OBS_TIME"))
    {
        // setInternalField (NMR_cpmg.cpp:47-75)
        const std::string mode = C.s("RESIDUAL_FIELD");
        if (mode == "uniform" || mode == "import") {
            field.resize((size_t)S.W * S.H * S.D);
            if (mode == "uniform") {
                const std::string d = C.s("GRADIENT_DIRECTION");
                const int32_t dir_code = (d == "0" || d == "x") ? 0 : (d == "1" || d == "y") ? 1 : 2;
                check(rwnmr_internal_field_uniform(S.solid.data(), S.W, S.H, S.D, S.resolution,
                                                   C.d("GRADIENT_VALUE"), dir_code, field.data()),
                      "internal field");
            } else {
                check(rwnmr_internal_field_read(C.s("PATH_TO_FIELD").c_str(), S.W, S.H, S.D, field.data()),
                      "internal field file (dimensions must match the image)");
            }
        }
    }

    // setName + createDirectoryForData (NMR_cpmg.cpp:96-110)
    void set_dir(const std::string &parent, const std::string &suffix)
    {
        const std::string name = parent + "/NMR_cpmg_t=" + trimmed(exposure, 2, exposure) + "ms" + suffix;
        dir = S.sim_dir + "/" + name;
        check(rwnmr_make_dirs(dir.c_str()), "creating " + dir);
    }

    // setNMRTimeFramework (NMR_cpmg.cpp:112-147)
    void time_framework(bool map)
    {
        S.time_framework(exposure);
        times.assign(1, 0.0);
        const double step_t = (double)S.steps_per_echo * S.dt();
        double t = 0.0;
        for (uint32_t e = 0; e < S.echoes; ++e) {
            t += step_t;
            times.push_back(t);
        }
        if (map)
            S.map_walk();
        S.update_relaxivity();
    }

    // run_simulation (NMR_cpmg.cpp:149-175): image-based walk or histogram
    // propagation, bulk relaxation, normalisation
    void run_simulation()
    {
        const std::string method = C.s("METHOD");
        amps.assign(S.echoes + 1, 0.0);
        if (method == "histogram") {
            std::vector<double> pen(S.hist.nb);
            const bool sig = S.rho_type == "sigmoid";
            if (sig && S.rho.size() < 8)
                die("RHO_TYPE sigmoid needs 8 RHO values");
            check(rwnmr_histogram_penalties(S.hist.bins.data(), S.hist.nb, S.rho.empty() ? 0.0 : S.rho[0],
                                            sig ? S.rho.data() : nullptr, S.voxel, S.D0, S.steps_per_echo,
                                            pen.data()),
                  "histogram penalties");
            std::vector<uint32_t> first((size_t)S.histograms), last((size_t)S.histograms);
            int32_t nh = 0;
            check(rwnmr_histogram_list(S.echoes, S.histograms, &nh, first.data(), last.data()), "histogram list");
            if (nh != S.hist.n_used)
                die("histogram method needs the map walk of this time framework");
            check(rwnmr_histogram_decay(S.hist.list_amps.data(), nh, S.hist.nb, first.data(), last.data(),
                                        pen.data(), amps.data()),
                  "histogram decay");
        } else {
            Walkers &w = S.w;
            w.reset_walk();
            w.reset_energy();
            rwnmr_walkers v = w.view(false);
            if (!field.empty()) {
                if (S.shift)
                    die("the internal-field CPMG does not support voxel division");
                std::vector<double> phase(w.n, 0.0);
                v.energy = nullptr;
                check(rwnmr_cpmg_field_run(S.ctx, S.img, &v, S.echoes, S.steps_per_echo, S.bc, field.data(), S.gamma,
                                           S.dt(), w.energy.data(), phase.data(), amps.data()),
                      "internal-field CPMG", S.ctx);
            } else {
                check(rwnmr_cpmg_run(S.ctx, S.img, &v, S.echoes, S.steps_per_echo, S.bc, S.shift, amps.data()),
                      "CPMG", S.ctx);
            }
        }
        if (C.b("APPLY_BULK")) {   // applyBulk (:346-359)
            const double bulk = -1.0 / S.bulk_time;
            for (size_t e = 0; e < amps.size(); ++e)
                amps[e] = exp(bulk * times[e]) * amps[e];
        }
        const double norm = 1.0 / amps[0];   // normalizeSignal (:369-385)
        for (double &a : amps)
            a = norm * a;
    }

    // applyLaplace (NMR_cpmg.cpp:387-453)
    void laplace()
    {
        std::vector<double> t(times.begin() + 1, times.end()), dec(amps.begin() + 1, amps.end());
        const int nt2 = (int)C.i("NUM_T2_BINS");
        const double noise_amp = C.d("NOISE_AMP");
        rwnmr_inverter_set_lambda_range(C.d("MIN_LAMBDA"), C.d("MAX_LAMBDA"));
        if (C.b("USE_T2_LOGSPACE")) {
            initialize_inverter((int)t.size(), t.data(), C.d("MIN_T2"), C.d("MAX_T2"), nt2, (int)C.i("NUM_LAMBDAS"),
                                (int)C.i("PRUNE_NUM"), noise_amp);
        } else {
            // linspace_double_vector (B/Laplace/src/vector_funcs.cpp:85-100)
            std::vector<double> bins(nt2);
            const double lo = C.d("MIN_T2"), step = (C.d("MAX_T2") - lo) / (nt2 - 1);
            double v = lo;
            for (int i = 0; i < nt2; ++i, v += step)
                bins[i] = v;
            initialize_inverter_with_t2_bins((int)t.size(), t.data(), nt2, bins.data(), (int)C.i("NUM_LAMBDAS"),
                                             (int)C.i("PRUNE_NUM"), noise_amp);
        }
        find_best_lambda((int)dec.size(), dec.data());
        t2_bins.assign(rwnmr_inverter_num_t2_bins(), 0.0);
        t2_amps.assign(t2_bins.size(), 0.0);
        get_used_t2_bins(t2_bins.data());
        if (!invert((int)dec.size(), dec.data(), t2_amps.data()))
            die("Laplace inversion failed");
        noise.assign(amps.size(), 0.0);
        if (noise_amp != 0.0) {
            std::vector<double> raw(dec.size());
            if (get_noise(raw.data()))
                std::copy(raw.begin(), raw.end(), noise.begin() + 1);
        }
    }

    void write(bool walkers, bool histogram, bool histogram_list, bool decay)
    {
        if (walkers)
            S.write_walkers(dir + "/cpmg_walkers.csv");
        if (histogram)
            S.write_histogram(dir + "/cpmg_histogram.csv", S.hist.bins, S.hist.amps);
        if (histogram_list)
            S.write_histogram_list(dir + "/cpmg_histList.csv");
        if (decay) {
            const std::string p = dir + "/cpmg_decay.csv", q = dir + "/cpmg_T2.csv";
            check(rwnmr_write_cpmg_decay(p.c_str(), (uint32_t)amps.size(), times.data(), amps.data(),
                                         noise.empty() ? nullptr : noise.data()),
                  "writing " + p);
            check(rwnmr_write_two_columns(q.c_str(), "NMRT2_bins,NMRT2_amps", (uint32_t)t2_bins.size(),
                                          t2_bins.data(), t2_amps.data()),
                  "writing " + q);
        }
    }
};

// rwnmrApp::CPMG + NMR_cpmg::run (NMR_cpmg.cpp:87-93)
void run_cpmg(Sim &S, const std::string &path)
{
    printf("-- CPMG to be executed:\n");
    std::unique_ptr<Conf> cc(cpmg_conf());
    const std::string def = S.config_root + "./default/cpmg.config";
    cc->read(def, false);
    if (path != "default" && path != def)
        cc->read(path, true);
    Cpmg cp(S, *cc);
    cp.set_dir("", "");
    cp.time_framework(true);
    cp.run_simulation();
    cp.laplace();
    cp.write(cc->b("SAVE_WALKERS"), cc->b("SAVE_HISTOGRAM"), cc->b("SAVE_HISTOGRAM_LIST"), cc->b("SAVE_DECAY"));
    printf("- cpmg executed succesfully (%u echoes)\n", S.echoes);
}

// ---------------------------------------------------------------- multitau

// rwnmrApp::MultiTau + NMR_multitau (NMR_multitau.cpp:23-220)
void run_multitau(Sim &S, const std::string &mt_path, const std::string &cpmg_path)
{
    printf("-- MultiTau to be executed:\n");
    // multitau_config reads the cpmg defaults file first (multitau_config.cpp:15-22)
    std::unique_ptr<Conf> mc(multitau_conf());
    const std::string cpmg_def = S.config_root + "./default/cpmg.config";
    const std::string mt_file = mt_path != "default" ? mt_path : S.config_root + "./default/multitau.config";
    mc->read(cpmg_def, false);
    if (mt_file != cpmg_def)
        mc->read(mt_file, mt_path != "default");
    std::unique_ptr<Conf> cc(cpmg_conf());
    cc->read(cpmg_def, false);
    if (cpmg_path != "default" && cpmg_path != cpmg_def)
        cc->read(cpmg_path, true);

    const std::string scale = mc->s("TAU_SCALE");
    std::vector<double> values = mc->list("TAU_VALUES", false);
    rwnmr_multitau_params mp;
    memset(&mp, 0, sizeof(mp));
    mp.tau_min = mc->d("TAU_MIN");
    mp.tau_max = mc->d("TAU_MAX");
    mp.tau_scale = scale == "manual" ? 2 : scale == "log" ? 1 : 0;
    mp.tau_points = mp.tau_scale == 2 ? (uint32_t)values.size() : (uint32_t)mc->i("TAU_POINTS");
    mp.tau_values = values.data();
    mp.complete_decay = mc->b("COMPLETE_DECAY");

    // setName (NMR_multitau.cpp:47-68)
    const std::string tmin = mp.tau_scale == 2 ? std::to_string(values.empty() ? 0.0 : values.front())
                                                : std::to_string(mp.tau_min);
    const std::string tmax = mp.tau_scale == 2 ? std::to_string(values.empty() ? 0.0 : values.back())
                                                : std::to_string(mp.tau_max);
    const std::string name = "/NMR_multitau_min=" + tmin.substr(0, std::to_string(mp.tau_min).find('.') + 3) +
                             "ms_max=" + tmax.substr(0, std::to_string(mp.tau_max).find('.') + 3) + "ms_pts=" +
                             std::to_string(mp.tau_points) + "_scale=" + scale;
    const std::string dir = S.sim_dir + "/" + name;
    check(rwnmr_make_dirs(dir.c_str()), "creating " + dir);

    std::vector<uint32_t> steps(mp.tau_points);
    std::vector<double> te(mp.tau_points), signal;
    check(rwnmr_multitau_sequence(&mp, S.dt(), steps.data(), te.data()), "multitau sequence");
    Cpmg cp(S, *cc);
    const double obs = cp.exposure;
    for (uint32_t i = 0; i < mp.tau_points; ++i) {
        // setExposureTime + setCPMG (NMR_multitau.cpp:112-131)
        S.steps_per_echo = steps[i];
        cp.exposure = mp.complete_decay ? obs : te[i];
        cp.time_framework(false);
        const double tei = steps[i] * S.dt();
        cp.set_dir(name, "_te=" + trimmed(tei, 3, tei));
        cp.run_simulation();   // runCPMG (:133-150)
        if (cp.amps.size() > 1)
            signal.push_back(cp.amps[1]);
        if (mp.complete_decay) {
            cp.laplace();
            cp.write(mc->b("SAVE_WALKERS"), mc->b("SAVE_HISTOGRAM"), mc->b("SAVE_HISTOGRAM_LIST"),
                     mc->b("SAVE_DECAY"));
        }
    }
    if (mc->b("SAVE_DECAY")) {   // writeDecay (:221-246)
        const std::string p = dir + "/multitau_decay.csv";
        const uint32_t n = signal.size() == te.size() ? (uint32_t)te.size() : 0u;
        check(rwnmr_write_two_columns(p.c_str(), "echo_time,signal", n, te.data(), signal.data()), "writing " + p);
    }
    printf("- multitau sequence executed succesfully (%u echo times)\n", mp.tau_points);
}

// ---------------------------------------------------------------- PFGSE

double mean_of(const std::vector<double> &v)
{
    double s = 0;
    for (double x : v)
        s += x;
    return s / (double)v.size();
}

double stdev_of(const std::vector<double> &v, double m)
{
    double s = 0.0;
    for (double x : v)
        s += (x - m) * (x - m);
    return sqrt(s / (double)v.size());
}

// NMR_PFGSE::getNormalDistributionSamples (NMR_pfgse.cpp:1893-1905): a new
// std::normal_distribution on the simulation generator per call
std::vector<double> normal_samples(std::mt19937 &g, double sd, uint32_t n)
{
    std::normal_distribution<double> dist(0.0, sd);
    std::vector<double> out(n);
    for (uint32_t i = 0; i < n; ++i)
        out[i] = dist(g);
    return out;
}

// rwnmrApp::PFGSE + NMR_PFGSE::run (NMR_pfgse.cpp:30-131)
void run_pfgse(Sim &S, const std::string &path)
{
    printf("-- PFGSE to be executed:\n");
    std::unique_ptr<Conf> pc(pfgse_conf());
    const std::string def = S.config_root + "./default/pfgse.config";
    pc->read(def, false);
    if (path != "default" && path != def)
        pc->read(path, true);
    Walkers &w = S.w;

    std::vector<double> gmax = pc->list("MAX_GRADIENT", false);
    if (gmax.size() < 3)
        gmax = {3000.0, 0.0, 0.0};
    const uint32_t points = (uint32_t)pc->i("GRADIENT_SAMPLES");
    const double pulse = pc->d("PULSE_WIDTH");
    const bool apply_bulk = pc->b("APPLY_BULK");
    double noise_amp = std::max(0.0, pc->d("NOISE_AMP"));
    const double target_snr = pc->d("TARGET_SNR") > 0.0 ? pc->d("TARGET_SNR") : -1.0;
    const std::string thr_type = pc->s("THRESHOLD_TYPE");
    const double thr_value = pc->d("THRESHOLD_VALUE");
    uint32_t thr_window = (uint32_t)pc->i("THRESHOLD_WINDOW");
    const bool allow_sampling = pc->b("ALLOW_WALKER_SAMPLING");
    const bool absorption = pc->b("APPLY_ABSORPTION");
    if (points < 2)
        die("GRADIENT_SAMPLES must be at least 2");
    if (S.shift)
        die("the PFGSE sequence does not support voxel division");

    // createTimeSamples (pfgse_config.cpp:206-227)
    std::vector<double> times;
    const std::string seq = pc->s("TIME_SEQ");
    if (seq == "manual") {
        times = pc->list("TIME_VALUES", false);
        std::sort(times.begin(), times.end());
    } else {
        const uint32_t nt = (uint32_t)pc->i("TIME_SAMPLES");
        const double a = pc->d("TIME_MIN"), b = pc->d("TIME_MAX"), step = (b - a) / ((double)nt - 1.0);
        for (uint32_t i = 0; i < nt; ++i)
            times.push_back(seq == "log" ? pow(10.0, a + step * i) : a + step * i);
    }
    if (times.empty())
        die("no PFGSE diffusion times");

    // gradients and wave vectors (setGradientVector / setVectorK, :272-313)
    std::vector<double> gx(points), gy(points), gz(points), gnorm(points), kx(points), ky(points), kz(points),
        knorm(points);
    rwnmr_pfgse_gradients(gmax[0], gmax[1], gmax[2], points, gx.data(), gy.data(), gz.data());
    for (uint32_t k = 0; k < points; ++k) {
        gnorm[k] = sqrt(gx[k] * gx[k] + gy[k] * gy[k] + gz[k] * gz[k]);
        kx[k] = rwnmr_pfgse_wave_vector(gx[k], pulse, S.gamma);
        ky[k] = rwnmr_pfgse_wave_vector(gy[k], pulse, S.gamma);
        kz[k] = rwnmr_pfgse_wave_vector(gz[k], pulse, S.gamma);
        knorm[k] = sqrt(kx[k] * kx[k] + ky[k] * ky[k] + kz[k] * kz[k]);
    }
    uint32_t adjust = points > 1 ? points : 0;   // setThresholdFromSamples(points) in the constructor

    // createDirectoryForData (:247-254)
    const std::string dir = S.sim_dir + "//NMR_pfgse";
    check(rwnmr_make_dirs((dir + "/timesamples").c_str()), "creating " + dir);

    // correctExposureTimes (:216-240): the inspection-length scale
    if (pc->b("APPLY_SCALE_FACTOR")) {
        const double L = pc->d("INSPECTION_LENGTH");
        const double f = (L * L) / S.D0;
        for (double &t : times)
            t *= f;
    }
    // runInitialMapSimulation (:322-348) and resetNMR (:145-182)
    S.time_framework(times.back());
    S.map_walk();
    S.update_relaxivity();
    w.reset_walk();
    w.reset_energy();
    std::fill(w.collisions.begin(), w.collisions.end(), 0u);

    const bool save = pc->b("SAVE_PFGSE");
    const std::string results = dir + "/PFGSE_results.csv";
    if (save) {   // presave (:1315-1324)
        check(rwnmr_write_pfgse_results_header(results.c_str()), "writing " + results);
        const std::string pp = dir + "/PFGSE_parameters.txt", gp = dir + "/PFGSE_gradient.csv";
        check(rwnmr_write_pfgse_parameters(pp.c_str(), S.D0, pulse, S.gamma, gmax[0], gmax[1], gmax[2], points,
                                           times.data(), (uint32_t)times.size()),
              "writing " + pp);
        check(rwnmr_write_pfgse_gradient(gp.c_str(), points, gx.data(), gy.data(), gz.data(), kx.data(), ky.data(),
                                         kz.data()),
              "writing " + gp);
    }

    const bool sampled = S.samples > 1 && allow_sampling;
    const uint64_t per_sample = w.n / S.samples;
    uint64_t taken = 0;
    double D_sat_std = 0.0, D_msd_std = 0.0;
    double msd_std[3] = {0, 0, 0}, dmsd_std[3] = {0, 0, 0};
    for (uint32_t ti = 0; ti < times.size(); ++ti) {
        const double delta = times[ti];
        S.time_framework(delta);   // set() (:88-96)
        std::vector<double> rhs(points);
        for (uint32_t k = 0; k < points; ++k)
            rhs[k] = (-1.0) * knorm[k] * knorm[k] * (delta - (pulse / 3.0));   // computeRHS (:450-453)
        // simulation_cuda (NMR_pfgse_cuda.cpp:536-905): the walk goes on from
        // the previous diffusion time for the missing steps
        if (S.sim_steps < taken)
            die("PFGSE diffusion times must ascend");
        const uint32_t inc = (uint32_t)(S.sim_steps - taken);
        taken += inc;
        std::vector<double> Mkt(points), Ms(sampled ? (size_t)S.samples * points : 0);
        rwnmr_walkers v = w.view(true);
        if (sampled)
            check(rwnmr_pfgse_run_sampled(S.ctx, S.img, &v, w.x0.data(), w.y0.data(), w.z0.data(), &inc, 1, kx.data(),
                                          ky.data(), kz.data(), points, 0, S.voxel, S.bc, 0, nullptr, S.samples,
                                          Ms.data(), Mkt.data(), nullptr),
                  "PFGSE", S.ctx);
        else
            check(rwnmr_pfgse_run(S.ctx, S.img, &v, w.x0.data(), w.y0.data(), w.z0.data(), &inc, 1, kx.data(),
                                  ky.data(), kz.data(), points, 0, S.voxel, S.bc, S.shift, 0, nullptr, Mkt.data(),
                                  nullptr),
                  "PFGSE", S.ctx);

        // runSequence (:488-505): bulk, noise vector, noise
        const double bulk = exp((-1.0 / S.bulk_time) * delta);
        if (apply_bulk)
            for (double &m : Mkt)
                m *= bulk;
        double basis = (double)w.n;
        if (target_snr > 0.0) {   // computeTargetNoiseAmp (:548-562)
            double sss = 0.0, uss = 0.0;
            for (uint32_t k = 0; k < points; ++k) {
                sss += Mkt[k] * Mkt[k];
                uss += Mkt[k];
            }
            sss /= (double)points;
            uss /= (double)points;
            noise_amp = sqrt((sss - uss * uss) / target_snr);
            basis = 1.0;
        }
        std::vector<double> raw_noise = normal_samples(S.rng, basis * noise_amp, points);
        if (noise_amp > 0.0)
            for (uint32_t k = 0; k < points; ++k)
                Mkt[k] += raw_noise[k];

        // recoverDsat (:605-621)
        std::vector<double> Mkt_std(points, 0.0), lhs(points), lhs_std(points, 0.0);
        double D_sat = 0.0, D_err = 0.0;
        auto apply_threshold = [&](const std::vector<double> &L) {   // applyThreshold (:133-143)
            const uint32_t win = std::min(thr_window, points);
            if (thr_type == "lhs") {
                if (L.size() >= win && thr_value > 0.0 && thr_value < 1.0)
                    adjust = noise_amp == 0.0
                                 ? rwnmr_pfgse_threshold_lhs(L.data(), points, thr_value, win)
                                 : rwnmr_pfgse_threshold_lhs_window(L.data(), points, thr_value, win);
            } else {
                const int v = thr_type == "samples" ? (int)thr_value : (int)points;
                if (v <= (int)points && v > 1)
                    adjust = (uint32_t)v;
            }
        };
        if (sampled) {   // recoverDsatWithSampling (:876-1060)
            std::vector<double> snoise((size_t)S.samples * points, 0.0);
            if (noise_amp > 0.0) {   // computeSamplesNoise (:833-874)
                const double M0 = (double)w.n / (double)S.samples;
                for (uint32_t s = 0; s < S.samples; ++s) {
                    std::vector<double> nz = normal_samples(S.rng, M0 * noise_amp, points);
                    std::copy(nz.begin(), nz.end(), snoise.begin() + (size_t)s * points);
                }
            }
            for (size_t q = 0; q < Ms.size(); ++q) {
                if (apply_bulk)
                    Ms[q] *= bulk;
                if (noise_amp > 0.0)
                    Ms[q] += snoise[q];
            }
            std::vector<double> mean_lhs(points), Ds(S.samples);
            check(rwnmr_pfgse_sample_stats(Ms.data(), S.samples, points, rhs.data(), points, Mkt.data(),
                                           Mkt_std.data(), mean_lhs.data(), lhs_std.data(), nullptr, nullptr, nullptr,
                                           nullptr),
                  "sample statistics");
            lhs = mean_lhs;
            for (uint32_t k = 0; k < points; ++k) {
                std::vector<double> col(S.samples);
                for (uint32_t s = 0; s < S.samples; ++s)
                    col[s] = snoise[(size_t)s * points + k];
                raw_noise[k] = mean_of(col);
            }
            apply_threshold(lhs);
            if (adjust < 2)
                die("PFGSE fit needs at least 2 points");
            check(rwnmr_pfgse_sample_stats(Ms.data(), S.samples, points, rhs.data(), adjust, nullptr, nullptr,
                                           nullptr, nullptr, Ds.data(), &D_sat, &D_sat_std, &D_err),
                  "sample fits");
        } else {   // recoverDsatWithoutSampling (:641-692)
            const double M0 = Mkt[0];
            std::vector<double> nm(points);
            for (uint32_t k = 0; k < points; ++k)
                nm[k] = Mkt[k] / M0;
            for (uint32_t k = 0; k < points; ++k) {
                const double r = nm[k] / nm[0];   // computeLHS (:474-481)
                lhs[k] = r > 1.0e-06 ? log(r) : log(1.0e-06);
            }
            apply_threshold(lhs);
            if (adjust < 2)
                die("PFGSE fit needs at least 2 points");
            check(rwnmr_stejskal_tanner(rhs.data(), lhs.data(), adjust, &D_sat, &D_err), "Stejskal-Tanner fit");
        }

        // recoverDmsd (:623-639, 1098-1291): displacement statistics over
        // walker samples (one sample of all walkers without sampling)
        const uint32_t ns = allow_sampling ? S.samples : 1u;
        const uint64_t per = allow_sampling ? per_sample : w.n;
        std::vector<double> sD, sDx, sDy, sDz, sM, sMx, sMy, sMz;
        for (uint32_t s = 0; s < ns; ++s) {
            double sq = 0.0, nx = 0.0, ny = 0.0, nz = 0.0, alive = 0.0;
            for (uint64_t j = 0; j < per; ++j) {
                const uint64_t i = (uint64_t)s * per + j;
                double dx = S.voxel * ((double)w.x[i] - (double)w.x0[i]);
                double dy = S.voxel * ((double)w.y[i] - (double)w.y0[i]);
                double dz = S.voxel * ((double)w.z[i] - (double)w.z0[i]);
                dx *= dx;
                dy *= dy;
                dz *= dz;
                const double f = absorption ? w.energy[i] : 1.0;
                alive += f;
                nx += f * dx;
                ny += f * dy;
                nz += f * dz;
                sq += f * (dx + dy + dz);
            }
            sq /= alive;
            nx /= alive;
            ny /= alive;
            nz /= alive;
            sM.push_back(sq);
            sD.push_back(sq / (6.0 * delta));
            sMx.push_back(nx);
            sMy.push_back(ny);
            sMz.push_back(nz);
            sDx.push_back(nx / (2.0 * delta));
            sDy.push_back(ny / (2.0 * delta));
            sDz.push_back(nz / (2.0 * delta));
        }
        const double D_msd = mean_of(sD);
        const double msd[3] = {mean_of(sMx), mean_of(sMy), mean_of(sMz)};
        const double dmsd[3] = {mean_of(sDx), mean_of(sDy), mean_of(sDz)};
        if (allow_sampling) {
            D_msd_std = stdev_of(sD, D_msd);
            msd_std[0] = stdev_of(sMx, msd[0]);
            msd_std[1] = stdev_of(sMy, msd[1]);
            msd_std[2] = stdev_of(sMz, msd[2]);
            dmsd_std[0] = stdev_of(sDx, dmsd[0]);
            dmsd_std[1] = stdev_of(sDy, dmsd[1]);
            dmsd_std[2] = stdev_of(sDz, dmsd[2]);
        }
        printf("D(%g ms) {s&t} = %.10g, {msd} = %.10g\n", delta, D_sat, D_msd);

        // save (:1326-1357)
        const std::string idx = std::to_string(ti);
        if (save) {
            const double row[11] = {D_sat,   D_err,       D_sat_std, D_msd,   D_msd_std,  dmsd[0],
                                    dmsd_std[0], dmsd[1], dmsd_std[1], dmsd[2], dmsd_std[2]};
            check(rwnmr_append_pfgse_results(results.c_str(), delta, row, adjust), "writing " + results);
            const std::string ep = dir + "/timesamples/echoes_" + idx + ".csv";
            check(rwnmr_write_pfgse_echoes(ep.c_str(), points, gnorm.data(), Mkt.data(), raw_noise.data(),
                                           Mkt_std.data(), lhs.data(), lhs_std.data(), rhs.data()),
                  "writing " + ep);
            const std::string mp = dir + "/timesamples/msd_" + idx + ".csv";
            check(rwnmr_write_pfgse_msd(mp.c_str(), msd, msd_std, dmsd, dmsd_std), "writing " + mp);
        }
        if (pc->b("SAVE_WALKERS"))
            S.write_walkers(dir + "/timesamples/walkers_" + idx + ".csv");
        if (pc->b("SAVE_HISTOGRAM")) {   // updateHistogram with the current steps
            std::vector<double> hb(S.hist.nb), ha(S.hist.nb);
            check(rwnmr_collision_histogram(w.collisions.data(), w.n, S.sim_steps, S.histogram_size,
                                            S.hist_log ? 1 : 0, hb.data(), ha.data()),
                  "histogram");
            S.write_histogram(dir + "/timesamples/histogram_" + idx + ".csv", hb, ha);
        }
        if (pc->b("SAVE_HISTOGRAM_LIST"))
            S.write_histogram_list(dir + "/timesamples/histList_" + idx + ".csv");
    }
    printf("- pfgse sequence executed succesfully\n");
}

}  // namespace

int main(int argc, char **argv)
{
    Args args(argc, argv);
    Sim S;
    const char *root = getenv("RWNMR_ROOT");
    S.root = root ? root : "";
    if (!S.root.empty() && S.root.back() != '/')
        S.root += "/";
    S.config_root = S.root + "./config/";
    printf("--- Building RWNMR essentials\n");
    build_essentials(S, args);
    // rwnmrApp::exec (B/App/rwnmrApp.cpp:74-88)
    for (size_t c = 2; c < args.commands.size(); ++c) {
        const std::string &cmd = args.commands[c];
        if (cmd == "cpmg")
            run_cpmg(S, args.paths[c]);
        else if (cmd == "pfgse")
            run_pfgse(S, args.paths[c]);
        else if (cmd == "multitau")
            run_multitau(S, args.paths[c], args.paths[c + 1]);
        else if (cmd == "ga")
            printf("-- GA is under construction.\n");   // rwnmrApp::GA (:127-130)
    }
    rwnmr_image_destroy(S.img);
    rwnmr_ctx_destroy(S.ctx);
    return 0;
}
