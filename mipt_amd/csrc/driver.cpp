// Whole runs on the batched engine: mipt_hip_run_batch (draws, evolution and
// per-trajectory observables in one call) and the sim_tmi scan driver.
//
// The scan keeps the contract of the reference's online TMI simulator
// (cpp/apps/sim_tmi.cpp:11-90, cpp/include/mipt/tmi/runner.hpp:52-409,
// 429-491, 556-590): positional arguments
//   n periods realizations resolution p_min p_max circ_type [all_cycles] [output.csv]
// (output.csv may also come before all_cycles), a linspace p grid
// (cpp/include/mipt/types.hpp:23-38), closed brickwork rings of 2*periods
// layers, one TMI row per (p, realization, cycle) in p-major order, CSV
// header `p,tmi`, numbers in 17 significant digits
// (cpp/include/mipt/tmi/linalg.hpp:223-234), the default path
// csv/tmi/<tag>/tmi_<tag>_n_<n>_real_<R>_<p_min>_<p_max>_res_<res>.csv, and
// SIM_TMI_ENTROPY_ORDER=2 for Renyi-2 TMI (tag suffix _s2).  What differs:
// every trajectory of the scan runs in one batched launch sequence instead of
// serially; only circ_type 1 (Haar U(4)) is built, and it is the default; the
// seeds are trajectory_seed(MIPT_SEED, index) when MIPT_SEED is set (random
// otherwise, as the reference's random_device).  An existing output is
// resumed as the reference does (SIM_TMI_RESUME, default on: whole circuits
// kept, a trailing partial circuit cut off, the p grid checked row by row).
#include "mipt_hip.h"
#include "common.hpp"

#include <algorithm>
#include <array>
#include <cerrno>
#include <charconv>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <chrono>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <random>
#include <sstream>
#include <stdexcept>
#include <string>
#include <string_view>
#include <sys/stat.h>
#include <unistd.h>
#include <cctype>
#include <iterator>
#include <limits>
#include <thread>
#include <vector>

namespace
{

struct HipError : std::runtime_error
{
    using std::runtime_error::runtime_error;
};

void check(int status, const char *what)
{
    if (status != 0)
    {
        const char *detail = mipt_hip_last_error();
        throw HipError(std::string(what) + " failed (status " + std::to_string(status) +
                       "): " + ((detail && *detail) ? detail : "no detail"));
    }
}

// cpp/include/mipt/types.hpp:262-273
std::uint64_t splitmix64(std::uint64_t v)
{
    v += 0x9e3779b97f4a7c15ull;
    v = (v ^ (v >> 30)) * 0xbf58476d1ce4e5b9ull;
    v = (v ^ (v >> 27)) * 0x94d049bb133111ebull;
    return v ^ (v >> 31);
}

std::uint64_t trajectory_seed(std::uint64_t master, std::uint64_t index)
{
    return splitmix64(master + 0x2545f4914f6cdd1dull * (index + 1ull));
}

// cpp/include/mipt/types.hpp:23-38: the last point is p_max exactly
// MIPT_SEED: a decimal (or 0x-prefixed hexadecimal) unsigned 64-bit value.
// A malformed value is refused rather than read as a prefix or as 0, so a
// resumed scan cannot continue on another seed stream unnoticed.
bool seed_from_env(std::uint64_t &master)
{
    const char *e = std::getenv("MIPT_SEED");
    if (!e) return false;
    const char *s = e;
    while (*s == ' ' || *s == '\t') ++s;
    if (*s == '\0' || *s == '-' || *s == '+')
        throw std::invalid_argument(std::string("MIPT_SEED must be an unsigned 64-bit integer, got '") + e + "'.");
    errno = 0;
    char *end = nullptr;
    const unsigned long long v = std::strtoull(s, &end, 0);
    while (end && (*end == ' ' || *end == '\t')) ++end;
    if (errno == ERANGE || end == s || *end != '\0')
        throw std::invalid_argument(std::string("MIPT_SEED must be an unsigned 64-bit integer, got '") + e + "'.");
    master = static_cast<std::uint64_t>(v);
    return true;
}

std::vector<double> linspace(double first, double last, int count)
{
    std::vector<double> v;
    if (count == 1)
    {
        v.push_back(first);
        return v;
    }
    const double step = (last - first) / (count - 1);
    for (int i = 0; i < count; ++i) v.push_back(first + step * i);
    v.back() = last;
    return v;
}

// 17 significant digits, shortest "general" form
void append_double(std::string &s, double value)
{
    char buf[64];
    const auto r = std::to_chars(buf, buf + sizeof(buf), value, std::chars_format::general, 17);
    if (r.ec == std::errc{}) s.append(buf, r.ptr);
    else s += std::to_string(value);
}

// fixed 12 decimals with trailing zeros (and a bare point) removed
std::string number_for_filename(double value)
{
    if (std::abs(value) < 1.0e-15) value = 0.0;
    std::ostringstream out;
    out << std::fixed << std::setprecision(12) << value;
    std::string text = out.str();
    if (text.find('.') != std::string::npos)
    {
        while (!text.empty() && text.back() == '0') text.pop_back();
        if (!text.empty() && text.back() == '.') text.pop_back();
    }
    if (text == "-0") text = "0";
    return text;
}

bool is_bool01(std::string_view s) { return s == "0" || s == "1"; }

bool parse_bool01(std::string_view s, const char *name)
{
    if (!is_bool01(s)) throw std::invalid_argument(std::string(name) + " must be 0 or 1.");
    return s == "1";
}

void mkdirs_for(const std::string &path)
{
    for (std::size_t i = 1; i < path.size(); ++i)
        if (path[i] == '/')
        {
            const std::string dir = path.substr(0, i);
            if (::mkdir(dir.c_str(), 0755) != 0 && errno != EEXIST)
                throw std::runtime_error("Could not create directory: " + dir);
        }
}

void usage(const char *argv0)
{
    std::cerr << "Usage:\n  " << argv0
              << " [n = 10] [periods = 10] [realizations = 10] [resolution = 5] [p_min = 0.0] [p_max = 1.0]"
                 " [circ_type = 1] [all_cycles = 0] [output.csv = auto]\n\n"
              << "Arguments:\n  circ_type:\n    1 = random Haar U(4) brickwork circuit (the only one built here)\n"
              << "  all_cycles = 0 computes the original offset only.\n"
              << "  all_cycles = 1 computes N/4 cyclic offsets.\n"
              << "  For backward compatibility, output.csv may also be supplied before all_cycles.\n"
              << "  If output.csv is omitted, the default path is:\n"
              << "    csv/tmi/<tag>/tmi_<tag>_n_<n>_real_<realizations>_<p_min>_<p_max>_res_<resolution>.csv\n"
              << "  The tag is haar; all_cycles=1 appends c, SIM_TMI_ENTROPY_ORDER=2 appends _s2.\n\n"
              << "Output CSV columns:\n  p,tmi\n\n"
              << "Environment:\n  SIM_TMI_ENTROPY_ORDER=1|2  von Neumann (default) or Renyi-2 TMI\n"
              << "  MIPT_SEED=<u64>            reproducible trajectory seeds\n"
              << "  MIPT_HIP_STATE_GIB=<g>     device memory for states per batch (default 64)\n";
}

bool env_flag(const char *name, bool dflt)
{
    const char *e = std::getenv(name);
    if (!e || !*e) return dflt;
    std::string v(e);
    for (char &ch : v) ch = static_cast<char>(std::tolower(static_cast<unsigned char>(ch)));
    if (v == "0" || v == "false" || v == "no" || v == "off") return false;
    if (v == "1" || v == "true" || v == "yes" || v == "on") return true;
    throw std::invalid_argument(std::string(name) + " must be 0/1, true/false, yes/no or on/off.");
}

// Fields of a generated output name tmi_<tag>_n_<n>_real_<R>_<p_min>_<p_max>_res_<res>.csv
struct NameIdentity
{
    std::string tag, n, real, pmin, pmax, res;
};

bool parse_output_name(std::string path, NameIdentity &id)
{
    const std::size_t slash = path.find_last_of('/');
    if (slash != std::string::npos) path = path.substr(slash + 1);
    if (path.size() < 8 || path.compare(0, 4, "tmi_") != 0 || path.compare(path.size() - 4, 4, ".csv") != 0)
        return false;
    const std::string body = path.substr(4, path.size() - 8);
    const std::size_t r = body.rfind("_res_"), q = body.find("_real_"), nn = body.rfind("_n_", q);
    if (r == std::string::npos || q == std::string::npos || nn == std::string::npos || !(nn < q && q < r)) return false;
    id.tag = body.substr(0, nn);
    id.n = body.substr(nn + 3, q - nn - 3);
    id.res = body.substr(r + 5);
    const std::string mid = body.substr(q + 6, r - q - 6);  // <R>_<p_min>_<p_max>
    const std::size_t a = mid.find('_'), b = mid.find('_', a == std::string::npos ? a : a + 1);
    if (a == std::string::npos || b == std::string::npos || mid.find('_', b + 1) != std::string::npos) return false;
    id.real = mid.substr(0, a);
    id.pmin = mid.substr(a + 1, b - a - 1);
    id.pmax = mid.substr(b + 1);
    return !id.tag.empty() && !id.n.empty() && !id.real.empty() && !id.res.empty();
}

// Where an interrupted scan stopped (the contract of the reference's
// cpp/include/mipt/tmi/resume.hpp:43-50, 205-333): header `p,tmi`; rows grouped
// by p in grid order, the p field compared as text against the grid rendered
// in 17 significant digits; a group switches only when full
// (realizations * cycles rows); only whole circuits (cycles rows) are kept,
// the file is cut back to the last one.
struct Checkpoint
{
    bool name_verified = false;
    std::uint64_t rows = 0, dropped_rows = 0;
    std::uintmax_t kept_bytes = 0;
    std::size_t start_p = 0;
    int start_real = 0;
};

[[noreturn]] void refuse(const std::string &why) { throw std::runtime_error(why); }

Checkpoint load_checkpoint(const std::string &path, const std::string &expected_name, int realizations, int cycles,
                           const std::vector<std::string> &p_texts)
{
    Checkpoint cp;
    NameIdentity want, got;
    if (parse_output_name(expected_name, want) && parse_output_name(path, got))
    {
        std::string mism;
        auto cmp = [&](const char *what, const std::string &a, const std::string &b) {
            if (a != b) mism += std::string(mism.empty() ? "" : ", ") + what + "=" + b + " (this run: " + a + ")";
        };
        cmp("tag", want.tag, got.tag);
        cmp("n", want.n, got.n);
        cmp("realizations", want.real, got.real);
        cmp("p_min", want.pmin, got.pmin);
        cmp("p_max", want.pmax, got.pmax);
        cmp("resolution", want.res, got.res);
        if (!mism.empty())
            refuse(path + " was produced with " + mism + ". Move the existing file aside, or set SIM_TMI_RESUME=0 to "
                                                        "overwrite it.");
        cp.name_verified = true;
    }
    std::ifstream in(path, std::ios::binary);
    std::string content((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    std::vector<std::string> lines;
    std::size_t pos = 0;
    while (true)
    {
        const std::size_t e = content.find('\n', pos);
        if (e == std::string::npos) break;  // an unterminated last line is not kept
        lines.push_back(content.substr(pos, e - pos));
        pos = e + 1;
    }
    if (lines.empty()) return cp;
    if (lines.front() != "p,tmi") refuse(path + " does not start with the expected header \"p,tmi\"; found \"" + lines.front() + "\".");
    const std::uint64_t per_p = static_cast<std::uint64_t>(realizations) * cycles;
    std::size_t group = 0;
    std::uint64_t in_group = 0, total = 0;
    std::uintmax_t offset = lines.front().size() + 1, boundary = offset;
    for (std::size_t i = 1; i < lines.size(); ++i)
    {
        const std::string &row = lines[i];
        const std::size_t comma = row.find(',');
        if (comma == std::string::npos) refuse(path + " has a malformed row at line " + std::to_string(i + 1) + ".");
        const std::string pt = row.substr(0, comma);
        if (pt != p_texts[group])
        {
            if (in_group != per_p)
                refuse(path + " switches away from p=" + p_texts[group] + " after " + std::to_string(in_group) +
                       " rows, but a complete p group holds realizations*cycles = " + std::to_string(per_p) + ".");
            ++group;
            if (group >= p_texts.size() || pt != p_texts[group])
                refuse(path + " carries p=" + pt + " where this run's grid expects " +
                       (group < p_texts.size() ? p_texts[group] : std::string("no further values")) + ".");
            in_group = 0;
        }
        ++in_group;
        ++total;
        offset += row.size() + 1;
        if (in_group % static_cast<std::uint64_t>(cycles) == 0) boundary = offset;
    }
    if (in_group > per_p) refuse(path + " holds more rows for p=" + p_texts[group] + " than a complete group.");
    const std::uint64_t whole = in_group / cycles;
    cp.dropped_rows = in_group - whole * cycles;
    cp.rows = total - cp.dropped_rows;
    cp.kept_bytes = boundary;
    if (whole >= static_cast<std::uint64_t>(realizations))
    {
        cp.start_p = group + 1;
        cp.start_real = 0;
    }
    else
    {
        cp.start_p = group;
        cp.start_real = static_cast<int>(whole);
    }
    return cp;
}

int entropy_order_from_env()
{
    const char *e = std::getenv("SIM_TMI_ENTROPY_ORDER");
    if (!e || !*e || std::string_view(e) == "1") return 1;
    if (std::string_view(e) == "2") return 2;
    throw std::invalid_argument("SIM_TMI_ENTROPY_ORDER must be 1 or 2.");
}

int run_sim_tmi(int argc, char **argv)
{
    if (argc > 1 && (std::strcmp(argv[1], "--help") == 0 || std::strcmp(argv[1], "-h") == 0))
    {
        usage(argv[0]);
        return 0;
    }
    if (argc > 10)
    {
        std::cerr << "Too many arguments. Use --help for usage.\n";
        return 2;
    }
    const int n = argc > 1 ? std::stoi(argv[1]) : 10;
    const int periods = argc > 2 ? std::stoi(argv[2]) : 10;
    const int realizations = argc > 3 ? std::stoi(argv[3]) : 10;
    const int res = argc > 4 ? std::stoi(argv[4]) : 5;
    const double p_min = argc > 5 ? std::stod(argv[5]) : 0.0;
    const double p_max = argc > 6 ? std::stod(argv[6]) : 1.0;
    const int circ = argc > 7 ? std::stoi(argv[7]) : 1;
    bool all_cycles = false;
    std::string output;
    if (argc > 8)
    {
        if (is_bool01(argv[8]))
        {
            all_cycles = parse_bool01(argv[8], "all_cycles");
            if (argc > 9) output = argv[9];
        }
        else
        {
            output = argv[8];
            if (argc > 9) all_cycles = parse_bool01(argv[9], "all_cycles");
        }
    }
    // cpp/include/mipt/tmi/runner.hpp:52-76
    if (n < 4 || (n % 4) != 0) throw std::invalid_argument("Online TMI requires n >= 4 and n divisible by 4.");
    if (periods <= 0) throw std::invalid_argument("periods must be positive.");
    if (realizations <= 0) throw std::invalid_argument("realizations must be positive.");
    if (res <= 0) throw std::invalid_argument("resolution must be positive.");
    if (!(0.0 <= p_min && p_min <= 1.0 && 0.0 <= p_max && p_max <= 1.0))
        throw std::invalid_argument("p_min and p_max must lie in [0, 1].");
    if (circ != 1)
        throw std::invalid_argument("circ_type " + std::to_string(circ) +
                                    " is not built on this engine; use 1 (random Haar U(4)).");
    if (n > 26) throw std::invalid_argument("this engine holds states up to n = 26.");
    const int order = entropy_order_from_env();
    std::string tag = "haar";
    if (all_cycles) tag += 'c';
    if (order == 2) tag += "_s2";
    const std::string default_path = "csv/tmi/" + tag + "/tmi_" + tag + "_n_" + std::to_string(n) + "_real_" +
                                     std::to_string(realizations) + "_" + number_for_filename(p_min) + "_" +
                                     number_for_filename(p_max) + "_res_" + std::to_string(res) + ".csv";
    if (output.empty()) output = default_path;
    const bool resume = env_flag("SIM_TMI_RESUME", true);

    std::uint64_t master;
    if (!seed_from_env(master)) master = (static_cast<std::uint64_t>(std::random_device{}()) << 32) ^ std::random_device{}();

    const int cycles = all_cycles ? n / 4 : 1;
    const std::vector<double> ps = linspace(p_min, p_max, res);
    const std::uint64_t total_traj = static_cast<std::uint64_t>(res) * realizations;
    double budget_gib = 64.0;
    if (const char *e = std::getenv("MIPT_HIP_STATE_GIB")) budget_gib = std::max(0.001, std::atof(e));
    const double state_bytes = std::ldexp(8.0, n);
    const std::uint64_t per_batch =
        std::max<std::uint64_t>(1, static_cast<std::uint64_t>(budget_gib * 1073741824.0 / state_bytes));

    // resume: keep the whole circuits an earlier run of the same scan wrote
    std::vector<std::string> p_texts;
    for (double pv : ps)
    {
        std::string t;
        append_double(t, pv);
        p_texts.push_back(t);
    }
    Checkpoint cp;
    struct stat sb;
    const bool exists = ::stat(output.c_str(), &sb) == 0;
    if (resume && exists) cp = load_checkpoint(output, default_path, realizations, cycles, p_texts);
    const std::uint64_t first_traj = static_cast<std::uint64_t>(cp.start_p) * realizations + cp.start_real;
    if (cp.rows >= total_traj * cycles)
    {
        std::cerr << output << " already holds all " << total_traj * cycles
                  << " TMI row(s) for this scan; nothing to do.\nSet SIM_TMI_RESUME=0 to overwrite it instead.\n";
        return 0;
    }
    const bool resuming = cp.rows > 0;
    mkdirs_for(output);
    if (resuming && cp.kept_bytes < static_cast<std::uintmax_t>(sb.st_size) &&
        ::truncate(output.c_str(), static_cast<off_t>(cp.kept_bytes)) != 0)
        throw std::runtime_error("Could not trim the partial circuit off " + output);
    std::ofstream csv(output, std::ios::binary | (resuming ? std::ios::app : std::ios::trunc));
    if (!csv) throw std::runtime_error("Could not create output CSV: " + output);
    if (!resuming) csv << "p,tmi\n";
    if (resuming)
        std::cerr << "RESUMED from " << output << ": " << cp.rows << " TMI row(s) already recorded"
                  << (cp.dropped_rows ? "; dropped " + std::to_string(cp.dropped_rows) + " row(s) of a partial circuit" : "")
                  << ". Continuing at p[" << cp.start_p << "]=" << ps[cp.start_p] << ", realization " << cp.start_real
                  << '/' << realizations << ".\n"
                  << (cp.name_verified ? "" : "Warning: the output name does not follow the generated convention; only "
                                              "the p grid and the per-p row counts were verified.\n");
    else if (!resume)
        std::cerr << "Resume disabled by SIM_TMI_RESUME=0; any existing output was overwritten.\n";
    std::cerr << "Online TMI simulation (batched HIP engine)\n"
              << "n=" << n << ", periods=" << periods << ", realizations=" << realizations << ", resolution=" << res
              << ", p_min=" << p_min << ", p_max=" << p_max << ", block_qubits=" << n / 4
              << ", all_cycles=" << (all_cycles ? 1 : 0) << ", tmi_cycles_per_circuit=" << cycles
              << ", mode=Random Haar Unitary, entropy_order=" << order << ", output=" << output
              << ", master_seed=" << master << ", trajectories_per_batch=" << per_batch << '\n';

    const auto t0 = std::chrono::steady_clock::now();
    std::string buf;
    std::uint64_t done = 0;
    for (std::uint64_t first = first_traj; first < total_traj; first += per_batch)
    {
        const int count = static_cast<int>(std::min<std::uint64_t>(per_batch, total_traj - first));
        std::vector<mipt_hip_traj_spec> specs(count);
        for (int i = 0; i < count; ++i)
        {
            const std::uint64_t idx = first + i;
            mipt_hip_traj_spec &s = specs[i];
            s.L = n;
            s.layers = 2 * periods;
            s.p = ps[idx / realizations];
            s.seed = trajectory_seed(master, idx);
            s.closed = 1;
            s.gate_family = 1;
            s.probe_site = -1;
            s.p_zero_layers = 0;
            s.shared_gates = -1;
            s.grid_x = 0;
        }
        std::vector<double> tmi(static_cast<std::size_t>(count) * cycles);
        std::vector<mipt_hip_traj_result> r(count);
        check(mipt_hip_run_batch(specs.data(), count, 0, cycles, order, tmi.data(), r.data(), nullptr),
              "run_batch");
        buf.clear();
        for (int i = 0; i < count; ++i)
            for (int c = 0; c < cycles; ++c)
            {
                append_double(buf, specs[i].p);
                buf.push_back(',');
                append_double(buf, tmi[static_cast<std::size_t>(i) * cycles + c]);
                buf.push_back('\n');
            }
        csv.write(buf.data(), static_cast<std::streamsize>(buf.size()));
        csv.flush();
        if (!csv) throw std::runtime_error("Failed while writing output CSV.");
        done += count;
        const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::cerr << "processed TMI " << (first_traj + done) * cycles << '/' << total_traj * cycles << " | avg " << std::fixed
                  << std::setprecision(2) << (done * cycles) / std::max(1e-9, el) << "/s" << std::defaultfloat
                  << '\n';
    }
    return 0;
}

// ------------------------------------------------------------------ mipt
// The three-site RDM record of the reference's mipt.exe
// (cpp/apps/mipt.cpp:79-197, 309-400; cpp/include/mipt/rdm.hpp:92-125;
// cpp/include/mipt/density.hpp:10-48): positional arguments
//   [d = 1] [n] [periods] [realizations] [resolution] [p_min] [p_max] [circ_type]
// info.csv plus rho3.bin (d = 1: the n cyclic triples (i, i+1, i+2) of a ring)
// or rho3_2d.bin (d = 2: an n x n grid, row triples then column triples,
// periodic), MIPTRHO2 version 2, one record per trajectory in (p, realization)
// order with the trace-normalised 8 x 8 matrices.  What differs: circ_type 1
// (Haar U(4)) only, also for d = 2 (the four-colour grid layering with Haar
// bonds, rows closed when n is even); 2 * periods layers; seeds as sim_tmi.
template <typename T>
void put(std::ofstream &f, const T &v)
{
    f.write(reinterpret_cast<const char *>(&v), sizeof(T));
}

int run_mipt(int argc, char **argv)
{
    if (argc > 1 && (std::strcmp(argv[1], "--help") == 0 || std::strcmp(argv[1], "-h") == 0))
    {
        std::cout << "Usage: " << argv[0]
                  << " [d = 1] [n = 10] [periods = 10] [realizations = 10] [resolution = 5] [p_min = 0.0] [p_max = 1.0]"
                     " [circ_type = 1]\n"
                  << "  d = 1: an n-site ring, the n cyclic three-site RDMs per trajectory into rho3.bin\n"
                  << "  d = 2: an n x n grid, 2 n^2 three-site RDMs (rows, then columns) into rho3_2d.bin\n"
                  << "  circ_type: 1 = random Haar U(4) (the only one built here)\n";
        return 0;
    }
    if (argc > 9)
    {
        std::cerr << "Too many arguments. Use --help for usage.\n";
        return 2;
    }
    const int d = argc > 1 ? std::stoi(argv[1]) : 1;
    const int n = argc > 2 ? std::stoi(argv[2]) : (d == 1 ? 10 : 4);
    const int periods = argc > 3 ? std::stoi(argv[3]) : (d == 1 ? 10 : 32);
    const int realizations = argc > 4 ? std::stoi(argv[4]) : 10;
    const int res = argc > 5 ? std::stoi(argv[5]) : 5;
    const double p_min = argc > 6 ? std::stod(argv[6]) : 0.0;
    const double p_max = argc > 7 ? std::stod(argv[7]) : 1.0;
    const int circ = argc > 8 ? std::stoi(argv[8]) : 1;
    // cpp/apps/mipt.cpp:342-366
    if (n <= 0) throw std::invalid_argument("n must be positive.");
    if (d != 1 && d != 2) throw std::invalid_argument("d must be 1 or 2.");
    if (res <= 0) throw std::invalid_argument("resolution must be positive.");
    if (realizations <= 0) throw std::invalid_argument("realizations must be positive.");
    if (periods < 0) throw std::invalid_argument("periods must be non-negative.");
    if (p_min < 0.0 || p_max > 1.0 || p_min > p_max) throw std::invalid_argument("Require 0 <= p_min <= p_max <= 1.");
    if (circ != 1)
        throw std::invalid_argument("circ_type " + std::to_string(circ) +
                                    " is not built on this engine; use 1 (random Haar U(4)).");
    const int L = d == 1 ? n : n * n;
    if (n < 3) throw std::invalid_argument("Saving three-qubit reduced density matrices requires n >= 3.");
    if (d == 1 && (n % 2) != 0) throw std::invalid_argument("Random Haar Unitary requires an even n for periodic brickwork geometry.");
    if (L > 26) throw std::invalid_argument("this engine holds states up to 26 qubits.");
    std::vector<int> subs;
    if (d == 1)
        for (int i = 0; i < n; ++i) subs.insert(subs.end(), {i, (i + 1) % n, (i + 2) % n});
    else
    {
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) subs.insert(subs.end(), {i * n + j, i * n + (j + 1) % n, i * n + (j + 2) % n});
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j)
                subs.insert(subs.end(), {j * n + i, ((j + 1) % n) * n + i, ((j + 2) % n) * n + i});
    }
    const int nsub = static_cast<int>(subs.size() / 3);
    std::uint64_t master;
    if (!seed_from_env(master)) master = (static_cast<std::uint64_t>(std::random_device{}()) << 32) ^ std::random_device{}();
    const std::vector<double> ps = linspace(p_min, p_max, res);
    const std::uint64_t total = static_cast<std::uint64_t>(res) * realizations;
    {
        std::ofstream info("info.csv");
        info << "n,periods,realizations,resolution,p_min,p_max,circ_type,circuit_name,fermion_trace,cudaq_backend\n";
        info << n << "," << periods << "," << realizations << "," << res << "," << p_min << "," << p_max << "," << circ
             << ",Random Haar Unitary,0,hip-gfx950\n";
        if (!info) throw std::runtime_error("Could not write info.csv");
    }
    const std::string path = d == 1 ? "rho3.bin" : "rho3_2d.bin";
    std::ofstream f(path, std::ios::binary | std::ios::trunc);
    if (!f) throw std::runtime_error("Could not create density-matrix file: " + path);
    const char magic[8] = {'M', 'I', 'P', 'T', 'R', 'H', 'O', '2'};
    f.write(magic, 8);
    put<std::uint32_t>(f, 0x01020304u);                 // endian marker
    put<std::uint32_t>(f, 2u);                          // version
    put<std::uint32_t>(f, static_cast<std::uint32_t>(d));
    put<std::uint32_t>(f, static_cast<std::uint32_t>(L));  // total qubits
    put<std::uint32_t>(f, 3u);                          // kept qubits
    put<std::uint32_t>(f, 8u);                          // matrix dimension
    put<std::uint32_t>(f, static_cast<std::uint32_t>(periods));
    put<std::uint32_t>(f, static_cast<std::uint32_t>(realizations));
    put<std::uint32_t>(f, static_cast<std::uint32_t>(res));
    put<std::uint32_t>(f, static_cast<std::uint32_t>(n));  // grid_x
    put<std::uint32_t>(f, static_cast<std::uint32_t>(d == 1 ? 1 : n));
    put<std::uint32_t>(f, static_cast<std::uint32_t>(nsub));
    put<std::uint64_t>(f, total);                       // records = trajectories
    put<double>(f, p_min);
    put<double>(f, p_max);
    for (int q : subs) put<std::uint32_t>(f, static_cast<std::uint32_t>(q));
    std::cout << "Running " << (d == 1 ? "1D " : "2D ") << L << "-qubit Haar simulation of " << total << " circuits.\n";
    double budget_gib = 64.0;
    if (const char *e = std::getenv("MIPT_HIP_STATE_GIB")) budget_gib = std::max(0.001, std::atof(e));
    const std::uint64_t per_batch =
        std::max<std::uint64_t>(1, static_cast<std::uint64_t>(budget_gib * 1073741824.0 / std::ldexp(8.0, L)));
    std::vector<double> rho;
    for (std::uint64_t first = 0; first < total; first += per_batch)
    {
        const int count = static_cast<int>(std::min<std::uint64_t>(per_batch, total - first));
        std::vector<mipt_hip_traj_spec> specs(count);
        for (int i = 0; i < count; ++i)
        {
            const std::uint64_t idx = first + i;
            specs[i] = mipt_hip_traj_spec{L, 2 * periods, ps[idx / realizations], trajectory_seed(master, idx), 1, 1, -1,
                                          0, -1, d == 1 ? 0 : n};
        }
        mipt_hip_batch *b = nullptr;
        const mipt_hip_batch_config cfg{0, 0, 0, 0};
        check(mipt_hip_batch_create(&cfg, nullptr, &b), "batch_create");
        int st = mipt_hip_batch_setup(b, specs.data(), count);
        if (!st) st = mipt_hip_batch_generate(b);
        if (!st) st = mipt_hip_batch_evolve(b);
        if (!st) st = mipt_hip_batch_synchronize(b);
        rho.assign(static_cast<std::size_t>(count) * nsub * 128, 0.0);
        if (!st) st = mipt_hip_batch_rho3(b, 0, count, subs.data(), nsub, rho.data());
        if (!st)
        {
            // a zero-weight branch anywhere in the batch is an error, as in the engine's readouts
            std::vector<int> o(static_cast<std::size_t>(2 * periods) * L);
            std::vector<std::uint8_t> fl(o.size());
            for (int i = 0; i < count && !st; ++i) st = mipt_hip_batch_outcomes(b, i, o.data(), fl.data(), nullptr, nullptr);
        }
        (void)mipt_hip_batch_destroy(b);
        check(st, "mipt batch");
        for (int i = 0; i < count; ++i)
        {
            const std::uint64_t idx = first + i;
            put<std::uint32_t>(f, static_cast<std::uint32_t>(idx / realizations));
            put<std::uint32_t>(f, static_cast<std::uint32_t>(idx % realizations));
            put<double>(f, specs[i].p);
            f.write(reinterpret_cast<const char *>(rho.data() + static_cast<std::size_t>(i) * nsub * 128),
                    static_cast<std::streamsize>(nsub * 128 * sizeof(double)));
        }
        if (!f) throw std::runtime_error("Failed while writing density-matrix payload.");
    }
    f.close();
    std::cout << "Saved " << nsub << " three-qubit reduced density matrices per trajectory to " << path << ".\n";
    return 0;
}

// --------------------------------------------------------------- entropy
// entropy.exe (cpp/apps/entropy.cpp:16-80, cpp/include/mipt/entropy.hpp:400-425,
// 521-640, cpp/include/mipt/entropy_resume.hpp:52-56): `n periods p
// realizations circ_type output.csv`; per trajectory the translation-averaged
// S1 (MIPT_ENTROPY_S1, MIPT_ENTROPY_S1_MAX_LA) and S2 in nats for blocks
// L_A = 2..n/2, accumulated over realizations (Welford), one CSV row per
// L_A.  What differs: circ_type 1 (Haar) only; no resume; seeds as sim_tmi.
struct Welford
{
    std::uint64_t count = 0;
    double mean = 0.0, m2 = 0.0;
    void add(double v)
    {
        ++count;
        const double d = v - mean;
        mean += d / static_cast<double>(count);
        m2 += d * (v - mean);
    }
    double sd() const { return count > 1 ? std::sqrt(std::max(0.0, m2 / static_cast<double>(count - 1))) : 0.0; }
    double se() const { return count > 1 ? sd() / std::sqrt(static_cast<double>(count)) : 0.0; }
};

int run_entropy(int argc, char **argv)
{
    if (argc > 1 && (std::strcmp(argv[1], "--help") == 0 || std::strcmp(argv[1], "-h") == 0))
    {
        std::cout << "Usage: " << argv[0] << " N periods p realizations circ_type output.csv\n"
                  << "  translation-averaged S1 and S2 (nats) of cyclic blocks L_A = 2..N/2\n"
                  << "  circ_type: 1 = random Haar U(4) (the only one built here)\n"
                  << "  MIPT_ENTROPY_S1=0 skips S1; MIPT_ENTROPY_S1_MAX_LA=k caps its block size\n";
        return 0;
    }
    if (argc != 7)
    {
        std::cout << "Usage: " << argv[0] << " N periods p realizations circ_type output.csv\n";
        return 2;
    }
    const int n = std::stoi(argv[1]);
    const int periods = std::stoi(argv[2]);
    const double p = std::stod(argv[3]);
    const int realizations = std::stoi(argv[4]);
    const int circ = std::stoi(argv[5]);
    const std::string output = argv[6];
    if (n < 4) throw std::invalid_argument("N must be at least 4 for the requested cyclic entropy scan.");
    if (periods < 0) throw std::invalid_argument("periods must be non-negative.");
    if (p < 0.0 || p > 1.0) throw std::invalid_argument("p must be in [0,1].");
    if (realizations <= 0) throw std::invalid_argument("realizations must be positive.");
    if (circ != 1)
        throw std::invalid_argument("circ_type " + std::to_string(circ) +
                                    " is not built on this engine; use 1 (random Haar U(4)).");
    if (n % 2 != 0) throw std::invalid_argument("Random Haar Unitary requires an even N for periodic brickwork geometry.");
    if (n > 26) throw std::invalid_argument("this engine holds states up to N = 26.");
    const int max_kept = n / 2, min_kept = 2;
    int max_s1 = env_flag("MIPT_ENTROPY_S1", true) ? max_kept : 0;
    if (const char *e = std::getenv("MIPT_ENTROPY_S1_MAX_LA"))
    {
        const long cap = std::atol(e);
        if (cap < 0 || cap > 62) throw std::invalid_argument("MIPT_ENTROPY_S1_MAX_LA must lie in [0, 62].");
        if (cap > 0 && max_s1 > 0) max_s1 = std::min<int>(max_s1, static_cast<int>(cap));
    }
    std::uint64_t master;
    if (!seed_from_env(master)) master = (static_cast<std::uint64_t>(std::random_device{}()) << 32) ^ std::random_device{}();
    const int sizes = max_kept - min_kept + 1;
    std::vector<Welford> s1(sizes), s2(sizes);
    double budget_gib = 64.0;
    if (const char *e = std::getenv("MIPT_HIP_STATE_GIB")) budget_gib = std::max(0.001, std::atof(e));
    const int per_batch = static_cast<int>(std::max<double>(1.0, std::min<double>(realizations, budget_gib * 1073741824.0 / std::ldexp(8.0, n))));
    for (int first = 0; first < realizations; first += per_batch)
    {
        const int count = std::min(per_batch, realizations - first);
        std::vector<mipt_hip_traj_spec> specs(count);
        for (int i = 0; i < count; ++i)
            specs[i] = mipt_hip_traj_spec{n, 2 * periods, p, trajectory_seed(master, static_cast<std::uint64_t>(first + i)),
                                          1, 1, -1, 0, -1, 0};
        mipt_hip_batch *b = nullptr;
        const mipt_hip_batch_config cfg{0, 0, 0, 0};
        check(mipt_hip_batch_create(&cfg, nullptr, &b), "batch_create");
        std::vector<double> v1(static_cast<std::size_t>(count) * sizes), v2(v1.size());
        int st = mipt_hip_batch_setup(b, specs.data(), count);
        if (!st) st = mipt_hip_batch_generate(b);
        if (!st) st = mipt_hip_batch_evolve(b);
        if (!st) st = mipt_hip_batch_synchronize(b);
        if (!st) st = mipt_hip_batch_translation_entropies(b, min_kept, max_kept, max_s1, v1.data(), v2.data());
        (void)mipt_hip_batch_destroy(b);
        check(st, "entropy batch");
        for (int i = 0; i < count; ++i)
            for (int k = 0; k < sizes; ++k)
            {
                if (min_kept + k <= max_s1) s1[k].add(v1[static_cast<std::size_t>(i) * sizes + k]);
                s2[k].add(v2[static_cast<std::size_t>(i) * sizes + k]);
            }
        std::cerr << "processed " << first + count << '/' << realizations << " realizations\n";
    }
    mkdirs_for(output);
    std::ofstream csv(output, std::ios::out | std::ios::trunc);
    if (!csv) throw std::runtime_error("Could not open output CSV: " + output);
    csv << std::setprecision(17);
    csv << "N,periods,p,requested_realizations,completed_realizations,circ_type,circuit_name,fermionic_trace,L_A,x,ln_x,"
           "S1_mean,S1_stddev,S1_stderr,S2_mean,S2_stddev,S2_stderr,subsystems_per_realization,total_subsystems\n";
    const double pi = 3.141592653589793238462643383279502884;
    for (int k = 0; k < sizes; ++k)
    {
        const int la = min_kept + k;
        const double x = (static_cast<double>(n) / pi) * std::sin(pi * la / static_cast<double>(n));
        const bool has1 = s1[k].count > 0;
        const double nan = std::numeric_limits<double>::quiet_NaN();
        csv << n << ',' << periods << ',' << p << ',' << realizations << ',' << s2[k].count << ',' << circ << ','
            << "\"Random Haar Unitary\"" << ',' << 0 << ',' << la << ',' << x << ',' << std::log(x) << ','
            << (has1 ? s1[k].mean : nan) << ',' << (has1 ? s1[k].sd() : nan) << ',' << (has1 ? s1[k].se() : nan) << ','
            << s2[k].mean << ',' << s2[k].sd() << ',' << s2[k].se() << ',' << n << ','
            << static_cast<std::uint64_t>(n) * s2[k].count << '\n';
    }
    csv.flush();
    if (!csv) throw std::runtime_error("Failed while writing output CSV: " + output);
    return 0;
}

// ------------------------------------------------------------ mipt_probed
// cpp/apps/mipt_probed.cpp with cpp/include/mipt/probed/config.hpp:261-372
// (arguments, reference sites, encoder schedule), runner.hpp:258-307 (the
// sampling trajectory) and output.hpp:36-95, 386-476 (names and CSV
// columns): modes 0 and 1 with one, two or four references on the Haar ring.
// One reference: Bell pair at t = 0 and t_encode = 2N unmeasured layers
// (mode 0: readout times are elapsed after the encoder; mode 1: t = 4N
// absolute).  Two or four: 2N monitored layers, the references attached at
// t0 = 2N, absolute readout times.  Every trajectory is read at every
// readout time in order; means and standard errors per (p, t).
struct ProbeStats
{
    Welford sq, i2, i3, i4, neg, logneg;
};

std::string compact_count(std::uint64_t count)
{
    if (count >= 1000000 && count % 1000000 == 0) return std::to_string(count / 1000000) + "m";
    if (count >= 1000 && count % 1000 == 0) return std::to_string(count / 1000) + "k";
    return std::to_string(count);
}

int env_int(const char *name, int dflt, long lo, long hi)
{
    const char *e = std::getenv(name);
    if (!e || !*e) return dflt;
    char *end = nullptr;
    const long v = std::strtol(e, &end, 10);
    if (*end != '\0' || v < lo || v > hi)
        throw std::invalid_argument(std::string(name) + " must be an integer in [" + std::to_string(lo) + ", " +
                                    std::to_string(hi) + "].");
    return static_cast<int>(v);
}

void probed_usage(const char *prog)
{
    std::cout << "Usage: " << prog << " [probes] [N] [realizations] [circ_type] [probe_mode] [args]\n\n"
              << "probes     1, 2 or 4 reference qubits\n"
              << "N          monitored system sites (even; N + probes <= 27)\n"
              << "circ_type  1 = random Haar U(4) brickwork ring (the only one built here)\n"
              << "mode 0     [p] [t_min] [t_max]: readout times t_min..t_max at one p\n"
              << "           (one reference: elapsed after the 2N-layer encoder, MIPT_PROBED_T_ENCODE;\n"
              << "            two/four: absolute, t_min >= t0 = 2N)\n"
              << "mode 1     [p_min] [p_max] [p_res]: readout at t = 4N\n"
              << "mode 2     [t_min] [t_max] [t_res] [p_min] [p_max] [p_res]: the (p, t) grid of S_Q, references\n"
              << "           Bell-paired at t = 0\n"
              << "Reference sites: one at N-MIPT_PROBED_DISTANCE (mode 0, default 1) or N/2 (mode 1);\n"
              << "two at (N-1)/2, N-1; four at N/4-1, N/2-1, 3N/4-1, N-1.\n"
              << "Output: the generated probed_*.csv name, or MIPT_PROBED_INTERNAL_OUTPUT; seeds from MIPT_SEED.\n";
}

// Information measures of `count` k-reference density matrices ([2^k][2^k]
// complex, consecutive) into v [count][11], on up to 16 host threads (the
// eigensolves are independent); returns the first failing status.
int probe_information_all(const double *rho, int count, int k, std::vector<double> &v)
{
    const std::size_t stride = static_cast<std::size_t>(2) << (2 * k);
    v.assign(static_cast<std::size_t>(count) * 11, 0.0);
    const int nthreads = std::max(1, std::min({16, count, static_cast<int>(std::thread::hardware_concurrency())}));
    std::vector<int> status(nthreads, 0);
    // the library's error text is thread-local: each worker keeps its own copy
    std::vector<std::string> message(nthreads);
    std::vector<std::thread> pool;
    for (int w = 0; w < nthreads; ++w)
        pool.emplace_back([&, w]() {
            for (int i = w; i < count && status[w] == 0; i += nthreads)
            {
                status[w] = mipt_hip_probe_information(rho + i * stride, k, v.data() + static_cast<std::size_t>(i) * 11);
                if (status[w]) message[w] = mipt_hip_last_error();
            }
        });
    for (auto &t : pool) t.join();
    for (int w = 0; w < nthreads; ++w)
        if (status[w])
        {
            mipt_hip::set_error("%s", message[w].c_str());  // re-raised in the calling thread
            return status[w];
        }
    return 0;
}

// Mode 4 with three references (probed/geometry.hpp:30-69, config.hpp:430-470,
// runner.hpp:94-148, output.hpp:216-267): every triangle of ring sites,
// grouped by its sorted periodic side lengths, kept when its chord balance
// (shortest / longest chord) reaches B_min; per geometry `realizations`
// trajectories on an embedding drawn from the trajectory seed, the three
// references attached after t_eq = 2N monitored layers, read out at
// tau = 0..tau_max: tripartite information, mean pair MI, minimum cut
// negativity, joint and mean single purity.  The genuine multipartite
// negativity needs a conic (SDP) solver and is not evaluated: its columns
// hold NaN and zero counts.
int run_probed_triangles(int argc, char **argv, int n, int realizations, int circ)
{
    const double p = std::stod(argv[6]);
    const int tau_max = std::stoi(argv[7]);
    if (p < 0.0 || p > 1.0 || tau_max < 0 || tau_max > (1 << 20))
        throw std::invalid_argument("Mode 4 requires p in [0,1] and a nonnegative tau_max "
                                    "that leaves room for the default t_eq=2N.");
    double bmin = 0.5;
    if (argc == 9) bmin = std::stod(argv[8]);
    else if (const char *e = std::getenv("MIPT_PROBED_TRIANGLE_BALANCE"); e && *e) bmin = std::atof(e);
    if (!std::isfinite(bmin) || bmin < 0.0 || bmin > 1.0)
        throw std::invalid_argument("Three-probe mode 4 requires B_min in [0,1].");
    const int t_eq = env_int("MIPT_PROBED_T_EQ", 2 * n, 0, 1L << 20);
    const double pi = std::acos(-1.0);
    auto chord = [&](int d) { return static_cast<double>(n) / pi * std::sin(pi * d / static_cast<double>(n)); };
    auto pdist = [&](int a, int b) { return std::min(std::abs(a - b), n - std::abs(a - b)); };
    struct Geometry
    {
        std::array<int, 3> d;
        double balance;
        std::vector<std::array<int, 3>> emb;
    };
    std::vector<Geometry> geo;  // sorted by side lengths, as an ordered map would hold them
    for (int a = 0; a < n - 2; ++a)
        for (int b = a + 1; b < n - 1; ++b)
            for (int c = b + 1; c < n; ++c)
            {
                std::array<int, 3> sig{pdist(a, b), pdist(a, c), pdist(b, c)};
                std::sort(sig.begin(), sig.end());
                auto it = std::lower_bound(geo.begin(), geo.end(), sig,
                                           [](const Geometry &g, const std::array<int, 3> &x) { return g.d < x; });
                if (it == geo.end() || it->d != sig) it = geo.insert(it, Geometry{sig, 0.0, {}});
                it->emb.push_back({a, b, c});
            }
    std::vector<Geometry> kept;
    for (auto &g : geo)
    {
        const double c0 = chord(g.d[0]), c1 = chord(g.d[1]), c2 = chord(g.d[2]);
        g.balance = std::min({c0, c1, c2}) / std::max({c0, c1, c2});
        if (g.balance + 1.0e-12 >= bmin) kept.push_back(g);
    }
    if (kept.empty())
        throw std::invalid_argument("No three-probe triangle geometry satisfies B >= " + std::to_string(bmin) +
                                    "; lower the mode-4 B_min argument or MIPT_PROBED_TRIANGLE_BALANCE.");
    std::string output;
    if (const char *e = std::getenv("MIPT_PROBED_INTERNAL_OUTPUT"); e && *e) output = e;
    else
        output = "probed_3_haar_n_" + std::to_string(n) + "_reals_" +
                 compact_count(static_cast<std::uint64_t>(realizations)) + "_mode_4_p_" + number_for_filename(p) +
                 "_b_" + number_for_filename(bmin) + "_teq_" + std::to_string(t_eq) + "_tau_0_" +
                 std::to_string(tau_max) + ".csv";
    std::uint64_t master;
    if (!seed_from_env(master)) master = (static_cast<std::uint64_t>(std::random_device{}()) << 32) ^ std::random_device{}();
    struct TriStats
    {
        Welford tmi, avg_mi, min_neg, joint, single;
    };
    const int ng = static_cast<int>(kept.size());
    std::vector<TriStats> stats(static_cast<std::size_t>(ng) * (tau_max + 1));
    const std::uint64_t total = static_cast<std::uint64_t>(ng) * realizations;
    double budget_gib = 64.0;
    if (const char *e = std::getenv("MIPT_HIP_STATE_GIB")) budget_gib = std::max(0.001, std::atof(e));
    const std::uint64_t per_batch = static_cast<std::uint64_t>(std::max<double>(
        1.0, std::min<double>(static_cast<double>(total), budget_gib * 1073741824.0 / std::ldexp(8.0, n + 3))));
    std::vector<double> rho;
    for (std::uint64_t first = 0; first < total; first += per_batch)
    {
        const int count = static_cast<int>(std::min<std::uint64_t>(per_batch, total - first));
        std::vector<mipt_hip_traj_spec> specs(count);
        std::vector<int> gid(count);
        for (int i = 0; i < count; ++i)
        {
            const std::uint64_t idx = first + i;
            gid[i] = static_cast<int>(idx / realizations);
            mipt_hip_traj_spec &s = specs[i];
            s = mipt_hip_traj_spec{};
            s.L = n;
            s.layers = t_eq + tau_max + 1;  // the attach layer t_eq lies inside the drawn range
            s.p = p;
            s.seed = trajectory_seed(master, idx);
            s.closed = 1;
            s.gate_family = 1;
            s.probe_site = -1;
            s.shared_gates = -1;
            s.probe_count = 3;
            const auto &emb = kept[gid[i]].emb;
            const auto &e = emb[splitmix64(s.seed ^ 0x4f524947494eull) % emb.size()];
            for (int j = 0; j < 3; ++j) s.probe_sites[j] = e[j];
            s.attach_layer = t_eq;
        }
        mipt_hip_batch *b = nullptr;
        const mipt_hip_batch_config cfg{0, 0, 0, 0};
        check(mipt_hip_batch_create(&cfg, nullptr, &b), "batch_create");
        int st = mipt_hip_batch_setup(b, specs.data(), count);
        if (!st) st = mipt_hip_batch_generate(b);
        rho.resize(static_cast<std::size_t>(count) * 128);
        int cur = 0;
        for (int tau = 0; tau <= tau_max && st == 0; ++tau)
        {
            if (tau == 0)
            {
                // right after the attach: three maximally mixed, uncorrelated references
                for (int i = 0; i < count; ++i)
                {
                    TriStats &a = stats[static_cast<std::size_t>(gid[i]) * (tau_max + 1)];
                    a.tmi.add(0.0), a.avg_mi.add(0.0), a.min_neg.add(0.0), a.joint.add(0.125), a.single.add(0.5);
                }
                continue;
            }
            const int target = t_eq + tau;
            if (target > cur)
            {
                st = mipt_hip_batch_evolve_range(b, cur, target);
                cur = target;
            }
            std::vector<double> info;
            if (!st) st = mipt_hip_batch_probe_rdm(b, rho.data());
            if (!st) st = probe_information_all(rho.data(), count, 3, info);
            for (int i = 0; i < count && st == 0; ++i)
            {
                const double *v = info.data() + static_cast<std::size_t>(i) * 11;
                TriStats &a = stats[static_cast<std::size_t>(gid[i]) * (tau_max + 1) + tau];
                a.tmi.add(v[6]), a.avg_mi.add(v[7]), a.min_neg.add(v[8]), a.joint.add(v[9]), a.single.add(v[10]);
            }
        }
        (void)mipt_hip_batch_destroy(b);
        check(st, "probed mode 4 batch");
        std::cerr << first + count << '/' << total << " trajectories\n";
    }
    int gmn_stride = env_int("MIPT_PROBED_GMN_STRIDE", 1, 1, 1L << 30);
    int gmn_limit = env_int("MIPT_PROBED_GMN_SAMPLES_PER_BIN", 0, 0, realizations);
    mkdirs_for(output);
    std::ofstream csv(output, std::ios::out | std::ios::trunc);
    if (!csv) throw std::runtime_error("Could not open output CSV: " + output);
    csv << std::setprecision(17);
    csv << "N,circ_type,circuit_name,realizations,p,geometry_id,distance_1,distance_2,distance_3,chord_1,chord_2,"
           "chord_3,chord_geometric_mean,triangle_balance,triangle_balance_cutoff,embedding_count,t_eq,tau,"
           "tau_over_chord_geometric_mean,t_absolute,TMI_mean,TMI_stderr,average_MI_mean,average_MI_stderr,"
           "min_bipartite_negativity_mean,min_bipartite_negativity_stderr,gmn_type,GMN_mean,GMN_stderr,"
           "GMN_sample_count,GMN_requested_count,GMN_zero_prefilter_count,GMN_solver_failure_count,GMN_stride,"
           "GMN_samples_per_bin,joint_purity_mean,joint_purity_stderr,mean_single_purity_mean,"
           "mean_single_purity_stderr,realizations_per_geometry\n";
    for (int g = 0; g < ng; ++g)
    {
        const Geometry &G = kept[g];
        const double c0 = chord(G.d[0]), c1 = chord(G.d[1]), c2 = chord(G.d[2]);
        const double cg = std::cbrt(c0 * c1 * c2);
        for (int tau = 0; tau <= tau_max; ++tau)
        {
            const TriStats &a = stats[static_cast<std::size_t>(g) * (tau_max + 1) + tau];
            csv << n << ',' << circ << ",\"Random Haar Unitary\"," << realizations << ',' << p << ',' << g << ','
                << G.d[0] << ',' << G.d[1] << ',' << G.d[2] << ',' << c0 << ',' << c1 << ',' << c2 << ',' << cg << ','
                << G.balance << ',' << bmin << ',' << G.emb.size() << ',' << t_eq << ',' << tau << ',' << tau / cg
                << ',' << t_eq + tau << ',' << a.tmi.mean << ',' << a.tmi.se() << ',' << a.avg_mi.mean << ','
                << a.avg_mi.se() << ',' << a.min_neg.mean << ',' << a.min_neg.se() << ",GMN,nan,nan,0,0,0,0,"
                << gmn_stride << ',' << gmn_limit << ',' << a.joint.mean << ',' << a.joint.se() << ','
                << a.single.mean << ',' << a.single.se() << ',' << realizations << '\n';
        }
    }
    csv.flush();
    if (!csv) throw std::runtime_error("Failed while writing output CSV: " + output);
    std::cout << output << '\n';
    return 0;
}

// Mode 4 with two references (probed/config.hpp:430-470, runner.hpp:94-148,
// 318-344, output.hpp:188-214): for every separation d = 1..N/2,
// `realizations` trajectories with the references at (x0, x0 + d mod N)
// from a uniform origin x0, t_eq = 2N monitored layers (MIPT_PROBED_T_EQ),
// both attached at t_eq, read out at tau = 0..tau_max.  The origin of
// trajectory i is splitmix64(seed_i ^ "ORIGIN") mod N, reproducible from the
// master seed (the reference draws it from a nondeterministic stream).
int run_probed_mode4(int argc, char **argv, int probes, int n, int realizations, int circ)
{
    if (probes != 2 && probes != 3) throw std::invalid_argument("Mode 4 requires probes=2 or probes=3.");
    if (n < probes) throw std::invalid_argument("Mode 4 requires N >= probes.");
    if ((probes == 2 && argc != 8) || (probes == 3 && argc != 8 && argc != 9))
        throw std::invalid_argument("Mode 4 requires [p] [tau_max], with optional [B_min] for probes=3.");
    if (probes == 3) return run_probed_triangles(argc, argv, n, realizations, circ);
    const double p = std::stod(argv[6]);
    const int tau_max = std::stoi(argv[7]);
    if (p < 0.0 || p > 1.0 || tau_max < 0 || tau_max > (1 << 20))
        throw std::invalid_argument("Mode 4 requires p in [0,1] and a nonnegative tau_max "
                                    "that leaves room for the default t_eq=2N.");
    const int t_eq = env_int("MIPT_PROBED_T_EQ", 2 * n, 0, 1L << 20);
    const int nd = n / 2;
    std::string output;
    if (const char *e = std::getenv("MIPT_PROBED_INTERNAL_OUTPUT"); e && *e) output = e;
    else
        output = "probed_2_haar_n_" + std::to_string(n) + "_reals_" +
                 compact_count(static_cast<std::uint64_t>(realizations)) + "_mode_4_p_" + number_for_filename(p) +
                 "_teq_" + std::to_string(t_eq) + "_tau_0_" + std::to_string(tau_max) + ".csv";
    std::uint64_t master;
    if (!seed_from_env(master)) master = (static_cast<std::uint64_t>(std::random_device{}()) << 32) ^ std::random_device{}();
    struct PairStats
    {
        Welford i2, neg, logneg;
        std::uint64_t positive = 0;
    };
    std::vector<PairStats> stats(static_cast<std::size_t>(nd) * (tau_max + 1));
    const std::uint64_t total = static_cast<std::uint64_t>(nd) * realizations;
    double budget_gib = 64.0;
    if (const char *e = std::getenv("MIPT_HIP_STATE_GIB")) budget_gib = std::max(0.001, std::atof(e));
    const std::uint64_t per_batch = static_cast<std::uint64_t>(
        std::max<double>(1.0, std::min<double>(static_cast<double>(total), budget_gib * 1073741824.0 / std::ldexp(8.0, n + 2))));
    const int depth = t_eq + tau_max;
    std::vector<double> rho;
    for (std::uint64_t first = 0; first < total; first += per_batch)
    {
        const int count = static_cast<int>(std::min<std::uint64_t>(per_batch, total - first));
        std::vector<mipt_hip_traj_spec> specs(count);
        std::vector<int> dist(count);
        for (int i = 0; i < count; ++i)
        {
            const std::uint64_t idx = first + i;
            dist[i] = static_cast<int>(idx / realizations) + 1;
            mipt_hip_traj_spec &s = specs[i];
            s = mipt_hip_traj_spec{};
            s.L = n;
            s.layers = depth + 1;  // the attach layer t_eq lies inside the drawn range even for tau_max = 0
            s.p = p;
            s.seed = trajectory_seed(master, idx);
            s.closed = 1;
            s.gate_family = 1;
            s.probe_site = -1;
            s.shared_gates = -1;
            s.probe_count = 2;
            const int x0 = static_cast<int>(splitmix64(s.seed ^ 0x4f524947494eull) % static_cast<std::uint64_t>(n));
            s.probe_sites[0] = x0;
            s.probe_sites[1] = (x0 + dist[i]) % n;
            s.attach_layer = t_eq;
        }
        mipt_hip_batch *b = nullptr;
        const mipt_hip_batch_config cfg{0, 0, 0, 0};
        check(mipt_hip_batch_create(&cfg, nullptr, &b), "batch_create");
        int st = mipt_hip_batch_setup(b, specs.data(), count);
        if (!st) st = mipt_hip_batch_generate(b);
        rho.resize(static_cast<std::size_t>(count) * 32);
        int cur = 0;
        for (int tau = 0; tau <= tau_max && st == 0; ++tau)
        {
            const int target = t_eq + tau;
            if (tau == 0)
            {
                // right after the attach the references are maximally mixed and uncorrelated
                for (int i = 0; i < count; ++i)
                {
                    PairStats &a = stats[static_cast<std::size_t>(dist[i] - 1) * (tau_max + 1)];
                    a.i2.add(0.0), a.neg.add(0.0), a.logneg.add(0.0);
                }
                continue;
            }
            if (target > cur)
            {
                st = mipt_hip_batch_evolve_range(b, cur, target);
                cur = target;
            }
            std::vector<double> info;
            if (!st) st = mipt_hip_batch_probe_rdm(b, rho.data());
            if (!st) st = probe_information_all(rho.data(), count, 2, info);
            for (int i = 0; i < count && st == 0; ++i)
            {
                const double *v = info.data() + static_cast<std::size_t>(i) * 11;
                PairStats &a = stats[static_cast<std::size_t>(dist[i] - 1) * (tau_max + 1) + tau];
                a.i2.add(v[1]), a.neg.add(v[4]), a.logneg.add(v[5]);
                a.positive += v[4] > 1.0e-12;
            }
        }
        (void)mipt_hip_batch_destroy(b);
        check(st, "probed mode 4 batch");
        std::cerr << first + count << '/' << total << " trajectories\n";
    }
    mkdirs_for(output);
    std::ofstream csv(output, std::ios::out | std::ios::trunc);
    if (!csv) throw std::runtime_error("Could not open output CSV: " + output);
    csv << std::setprecision(17);
    csv << "N,circ_type,circuit_name,realizations,p,distance,chord_length,t_eq,tau,tau_over_chord,t_absolute,I_mean,"
           "I_stderr,negativity_mean,negativity_stderr,log_negativity_mean,log_negativity_stderr,"
           "negativity_positive_count,negativity_positive_fraction,origins_randomized\n";
    const double pi = std::acos(-1.0);
    for (int d = 1; d <= nd; ++d)
    {
        const double chord = static_cast<double>(n) / pi * std::sin(pi * d / static_cast<double>(n));
        for (int tau = 0; tau <= tau_max; ++tau)
        {
            const PairStats &a = stats[static_cast<std::size_t>(d - 1) * (tau_max + 1) + tau];
            csv << n << ',' << circ << ",\"Random Haar Unitary\"," << realizations << ',' << p << ',' << d << ','
                << chord << ',' << t_eq << ',' << tau << ',' << tau / chord << ',' << t_eq + tau << ',' << a.i2.mean
                << ',' << a.i2.se() << ',' << a.neg.mean << ',' << a.neg.se() << ',' << a.logneg.mean << ','
                << a.logneg.se() << ',' << a.positive << ','
                << static_cast<double>(a.positive) / static_cast<double>(a.neg.count) << ",1\n";
        }
    }
    csv.flush();
    if (!csv) throw std::runtime_error("Failed while writing output CSV: " + output);
    std::cout << output << '\n';
    return 0;
}

int run_probed(int argc, char **argv)
{
    if (argc > 1 && (std::strcmp(argv[1], "--help") == 0 || std::strcmp(argv[1], "-h") == 0))
    {
        probed_usage(argv[0]);
        return 0;
    }
    if (argc < 6) throw std::invalid_argument("Missing arguments. Use --help for usage.");
    const int probes = std::stoi(argv[1]);
    const int n = std::stoi(argv[2]);
    const int realizations = std::stoi(argv[3]);
    const int circ = std::stoi(argv[4]);
    const int mode = std::stoi(argv[5]);
    if (probes != 1 && probes != 2 && probes != 3 && probes != 4)
        throw std::invalid_argument("probes must be 1, 2, 3, or 4.");
    if (n <= 0 || realizations <= 0) throw std::invalid_argument("N and realizations must be positive.");
    if (circ != 1)
        throw std::invalid_argument("circ_type " + std::to_string(circ) +
                                    " is not built on this engine; use 1 (random Haar U(4)).");
    if (n % 2 != 0) throw std::invalid_argument("Random Haar Unitary requires an even N for periodic brickwork geometry.");
    if (n + probes > 27) throw std::invalid_argument("this engine holds states up to N + probes = 27 qubits.");
    if (mode < 0 || mode > 5) throw std::invalid_argument("probe_mode must be 0, 1, 2, 3, 4, or 5.");
    if (probes == 3 && mode != 4) throw std::invalid_argument("Three probes are currently supported only in mode 4.");
    if (mode == 4) return run_probed_mode4(argc, argv, probes, n, realizations, circ);
    if (mode > 3)
        throw std::invalid_argument("probe_mode " + std::to_string(mode) + " is not built on this engine; use 0 to 4.");
    if (mode == 3 && probes != 2) throw std::invalid_argument("Mode 3 requires probes=2.");
    if (mode == 3 && argc != 9) throw std::invalid_argument("Mode 3 requires [p] [delta_x] [delta_t].");
    if (mode < 2 && argc != 9)
        throw std::invalid_argument(mode == 0 ? "Mode 0 requires [p] [t_min] [t_max]." : "Mode 1 requires [p_min] [p_max] [p_res].");
    if (mode == 2 && argc != 12)
        throw std::invalid_argument("Mode 2 requires [t_min] [t_max] [t_res] [p_min] [p_max] [p_res].");
    std::vector<double> ps;
    std::vector<int> times;
    int t_encode = 0, origin = 0;
    // mode 2 and one reference: Bell pairs at t = 0; two/four in modes 0/1: attach at t0 = 2N;
    // mode 3: reference 0 at tau1 = 4N, reference 1 delta_t later (t0 = tau2)
    int delta_x = 0, delta_t = 0;
    if (mode == 3)
    {
        delta_x = std::stoi(argv[7]);
        delta_t = std::stoi(argv[8]);
    }
    const int t0 = mode == 3 ? 4 * n + delta_t : (probes == 1 || mode == 2) ? 0 : 2 * n;
    const int attach0 = mode == 3 ? 4 * n : t0;
    if (mode == 3)
    {
        const double p = std::stod(argv[6]);
        if (p < 0.0 || p > 1.0 || delta_x < 0 || delta_x > n / 2 || delta_t < 0)
            throw std::invalid_argument("Mode 3 requires p in [0,1], 0 <= delta_x <= N/2, and delta_t >= 0.");
        if (delta_t == 0 && delta_x != n / 2)
            throw std::invalid_argument("The mode-3 spatial branch requires even N, delta_x=N/2, delta_t=0.");
        if (delta_t != 0 && delta_x != 0)
            throw std::invalid_argument("The mode-3 temporal branch requires delta_x=0 and delta_t>0.");
        ps = {p};
        for (int tau = 0; tau <= 8 * n; ++tau) times.push_back(tau);
        origin = t0;
    }
    else if (mode == 2)
    {
        const int t_min = std::stoi(argv[6]), t_max = std::stoi(argv[7]), t_res = std::stoi(argv[8]);
        const double p_min = std::stod(argv[9]), p_max = std::stod(argv[10]);
        const int p_res = std::stoi(argv[11]);
        if (t_min < 0 || t_max < t_min || p_min < 0.0 || p_max > 1.0)
            throw std::invalid_argument("Require nonnegative ordered times and p endpoints in [0,1].");
        if (t_res <= 0 || p_res <= 0) throw std::invalid_argument("Grid resolution must be positive.");
        if (p_min > p_max) throw std::invalid_argument("Grid minimum must not exceed its maximum.");
        // integer times: rounded linspace, no duplicates (util/grid.hpp:45-64)
        for (double v : linspace(t_min, t_max, t_res))
        {
            const int r = static_cast<int>(std::llround(v));
            if (times.empty() || times.back() != r) times.push_back(r);
        }
        if (static_cast<int>(times.size()) != t_res)
            throw std::invalid_argument("t_res requests duplicate integer times; require t_res <= t_max-t_min+1.");
        ps = linspace(p_min, p_max, p_res);
    }
    else if (mode == 0)
    {
        const double p = std::stod(argv[6]);
        const int t_min = std::stoi(argv[7]), t_max = std::stoi(argv[8]);
        if (p < 0.0 || p > 1.0 || t_min < 0 || t_max < t_min)
            throw std::invalid_argument("Require p in [0,1] and 0 <= t_min <= t_max.");
        if (probes > 1 && t_min < t0) throw std::invalid_argument("Two/four-probe legacy mode requires t_min >= t0=2N.");
        if (probes == 1) origin = t_encode = env_int("MIPT_PROBED_T_ENCODE", 2 * n, 0, 1L << 20);
        ps = {p};
        for (int t = t_min; t <= t_max; ++t) times.push_back(t);
    }
    else
    {
        const double p_min = std::stod(argv[6]), p_max = std::stod(argv[7]);
        const int p_res = std::stoi(argv[8]);
        if (p_min < 0.0 || p_max > 1.0) throw std::invalid_argument("p endpoints must lie in [0,1].");
        if (p_res <= 0) throw std::invalid_argument("Grid resolution must be positive.");
        if (p_min > p_max) throw std::invalid_argument("Grid minimum must not exceed its maximum.");
        ps = linspace(p_min, p_max, p_res);
        times = {4 * n};
        if (probes == 1) t_encode = 2 * n;
    }
    std::vector<int> sites;
    if (probes == 1)
        sites = {mode == 0 ? n - env_int("MIPT_PROBED_DISTANCE", 1, 1, n) : n / 2};
    else if (mode == 3)
        sites = {0, delta_x % n};
    else if (probes == 2)
        sites = {(n - 1) / 2, n - 1};
    else
    {
        if (n % 4 != 0) throw std::invalid_argument("Four-probe protocols require N divisible by four.");
        sites = {n / 4 - 1, n / 2 - 1, 3 * n / 4 - 1, n - 1};
    }
    const int depth = origin + times.back();

    std::string output;
    if (const char *e = std::getenv("MIPT_PROBED_INTERNAL_OUTPUT"); e && *e) output = e;
    else
    {
        output = "probed_" + std::to_string(probes) + "_haar_n_" + std::to_string(n) + "_reals_" +
                 compact_count(static_cast<std::uint64_t>(realizations)) + "_mode_" + std::to_string(mode);
        output += ps.size() == 1 ? "_p_" + number_for_filename(ps.front())
                                 : "_p_" + number_for_filename(ps.front()) + "_" + number_for_filename(ps.back()) +
                                       "_res_" + std::to_string(ps.size());
        if (mode == 3)
            output += "_dx_" + std::to_string(delta_x) + "_dt_" + std::to_string(delta_t) + "_tau_" +
                      std::to_string(times.front()) + "_" + std::to_string(times.back());
        else
            output += "_t_" + std::to_string(times.front()) + "_" + std::to_string(times.back());
        if (mode == 2) output += "_res_" + std::to_string(times.size());
        output += ".csv";
    }
    std::uint64_t master;
    if (!seed_from_env(master)) master = (static_cast<std::uint64_t>(std::random_device{}()) << 32) ^ std::random_device{}();

    const int D = 1 << probes;
    double budget_gib = 64.0;
    if (const char *e = std::getenv("MIPT_HIP_STATE_GIB")) budget_gib = std::max(0.001, std::atof(e));
    const int per_batch = static_cast<int>(
        std::max<double>(1.0, std::min<double>(realizations, budget_gib * 1073741824.0 / std::ldexp(8.0, n + probes))));
    std::vector<ProbeStats> stats(ps.size() * times.size());
    for (std::size_t pi = 0; pi < ps.size(); ++pi)
    {
        for (int first = 0; first < realizations; first += per_batch)
        {
            const int count = std::min(per_batch, realizations - first);
            std::vector<mipt_hip_traj_spec> specs(count);
            for (int i = 0; i < count; ++i)
            {
                mipt_hip_traj_spec &s = specs[i];
                s = mipt_hip_traj_spec{};
                s.L = n;
                s.layers = depth + 1;  // the attach layer t0 lies inside the drawn range
                s.p = ps[pi];
                s.seed = trajectory_seed(master, pi * static_cast<std::uint64_t>(realizations) + first + i);
                s.closed = 1;
                s.gate_family = 1;
                s.probe_site = -1;
                s.p_zero_layers = probes == 1 ? std::min(t_encode, depth) : 0;
                s.shared_gates = -1;
                s.grid_x = 0;
                s.probe_count = probes;
                for (int j = 0; j < probes; ++j) s.probe_sites[j] = sites[j];
                s.attach_layer = attach0;
                s.attach_stagger = delta_t;
            }
            mipt_hip_batch *b = nullptr;
            const mipt_hip_batch_config cfg{0, 0, 0, 0};
            check(mipt_hip_batch_create(&cfg, nullptr, &b), "batch_create");
            int st = mipt_hip_batch_setup(b, specs.data(), count);
            if (!st) st = mipt_hip_batch_generate(b);
            std::vector<double> rho(static_cast<std::size_t>(count) * D * D * 2);
            int cur = 0;
            for (std::size_t ti = 0; ti < times.size() && st == 0; ++ti)
            {
                const int target = origin + times[ti];
                ProbeStats &acc = stats[pi * times.size() + ti];
                if (target == 0)
                {
                    // before any layer every reference is maximally entangled
                    // with its site (Bell pairs at t = 0): S = k ln 2, no information
                    for (int i = 0; i < count; ++i)
                    {
                        acc.sq.add(probes * std::log(2.0));
                        acc.i2.add(0.0), acc.i3.add(0.0), acc.i4.add(0.0), acc.neg.add(0.0), acc.logneg.add(0.0);
                    }
                    continue;
                }
                if (probes > 1 && target == t0)
                {
                    // read right after the attach (which the engine runs at the
                    // start of layer t0): each reset site forms a fresh Bell pair
                    // with its reference, so the references are exactly in the
                    // maximally mixed product state: S = k ln 2, no information
                    for (int i = 0; i < count; ++i)
                    {
                        acc.sq.add(probes * std::log(2.0));
                        acc.i2.add(0.0), acc.i3.add(0.0), acc.i4.add(0.0), acc.neg.add(0.0), acc.logneg.add(0.0);
                    }
                    continue;
                }
                if (target > cur)
                {
                    st = mipt_hip_batch_evolve_range(b, cur, target);
                    cur = target;
                }
                std::vector<double> info;
                if (!st) st = mipt_hip_batch_probe_rdm(b, rho.data());
                if (!st) st = probe_information_all(rho.data(), count, probes, info);
                for (int i = 0; i < count && st == 0; ++i)
                {
                    const double *v = info.data() + static_cast<std::size_t>(i) * 11;
                    acc.sq.add(v[0]), acc.i2.add(v[1]), acc.i3.add(v[2]), acc.i4.add(v[3]);
                    acc.neg.add(v[4]), acc.logneg.add(v[5]);
                }
            }
            (void)mipt_hip_batch_destroy(b);
            check(st, "probed batch");
            std::cerr << "p=" << ps[pi] << ": " << first + count << '/' << realizations << " realizations\n";
        }
    }
    mkdirs_for(output);
    std::ofstream csv(output, std::ios::out | std::ios::trunc);
    if (!csv) throw std::runtime_error("Could not open output CSV: " + output);
    csv << std::setprecision(17);
    const char *meta = "N,circ_type,circuit_name,realizations,p";
    if (mode == 3)
        csv << meta << ",delta_x,delta_t,tau1,tau2,tau,tau_over_L,t_absolute,I_mean,I_stderr,negativity_mean,"
                       "negativity_stderr,log_negativity_mean,log_negativity_stderr\n";
    else if (mode == 2) csv << meta << ",t,S_Q_mean,S_Q_stderr\n";
    else if (mode == 0 && probes == 1) csv << meta << ",t,S_mean,S_stderr\n";
    else if (mode == 0 && probes == 2)
        csv << meta << ",t,t_minus_t0,scaled_time,I_mean,I_stderr,negativity_mean,negativity_stderr,"
                       "log_negativity_mean,log_negativity_stderr\n";
    else if (mode == 0)
        csv << meta << ",t,t_minus_t0,scaled_time,I2_mean,I2_stderr,I3_mean,I3_stderr,I4_mean,I4_stderr\n";
    else
    {
        csv << meta << ",t,S_Q_mean,S_Q_stderr";
        if (probes >= 2) csv << ",I2_mean,I2_stderr";
        if (probes == 4) csv << ",I3_mean,I3_stderr,I4_mean,I4_stderr";
        csv << '\n';
    }
    for (std::size_t pi = 0; pi < ps.size(); ++pi)
        for (std::size_t ti = 0; ti < times.size(); ++ti)
        {
            const ProbeStats &a = stats[pi * times.size() + ti];
            const int t = times[ti];
            csv << n << ',' << circ << ",\"Random Haar Unitary\"," << realizations << ',' << ps[pi];
            if (mode != 3) csv << ',' << t;
            if (mode == 3)
                csv << ',' << delta_x << ',' << delta_t << ',' << 4 * n << ',' << t0 << ',' << t << ','
                    << static_cast<double>(t) / n << ',' << t0 + t << ',' << a.i2.mean << ',' << a.i2.se() << ','
                    << a.neg.mean << ',' << a.neg.se() << ',' << a.logneg.mean << ',' << a.logneg.se();
            else if (mode == 2 || (mode == 0 && probes == 1)) csv << ',' << a.sq.mean << ',' << a.sq.se();
            else if (mode == 0)
            {
                csv << ',' << t - t0 << ',' << static_cast<double>(t - t0) / n << ',' << a.i2.mean << ',' << a.i2.se();
                if (probes == 2)
                    csv << ',' << a.neg.mean << ',' << a.neg.se() << ',' << a.logneg.mean << ',' << a.logneg.se();
                else
                    csv << ',' << a.i3.mean << ',' << a.i3.se() << ',' << a.i4.mean << ',' << a.i4.se();
            }
            else
            {
                csv << ',' << a.sq.mean << ',' << a.sq.se();
                if (probes >= 2) csv << ',' << a.i2.mean << ',' << a.i2.se();
                if (probes == 4) csv << ',' << a.i3.mean << ',' << a.i3.se() << ',' << a.i4.mean << ',' << a.i4.se();
            }
            csv << '\n';
        }
    csv.flush();
    if (!csv) throw std::runtime_error("Failed while writing output CSV: " + output);
    std::cout << output << '\n';
    return 0;
}

// ------------------------------------------------------------ free_energy
// cpp/apps/free_energy.cpp with cpp/include/mipt/free_energy.hpp:127-246
// (initial states) and free_energy_resume.hpp:79-88, 152-167 (CSV headers,
// the half-chain entropy schedule): the measurement-record free energy on
// the Haar ring.  4N equilibration + 24N production layers per trajectory;
// F(t) = cumulative surprisal -ln P(outcome) of the measurements (nats, from
// the engine's conditional P(1) of every measured site), S(L/2) in bits at
// t = 0 and every stride-th t <= entropy_max_t; per-trajectory samples and
// per-t aggregates in the reference's two CSVs (no resume: existing files
// are refused unless MIPT_FREE_ENERGY_RESUME=0).
std::uint64_t fe_splitmix64(std::uint64_t v)
{
    v += 0x9e3779b97f4a7c15ull;
    v = (v ^ (v >> 30)) * 0xbf58476d1ce4e5b9ull;
    v = (v ^ (v >> 27)) * 0x94d049bb133111ebull;
    return v ^ (v >> 31);
}

// The initial state of trajectory `trajectory_seed` (free_energy.hpp:127-222
// draw order, std::mt19937_64 keyed as free_energy.cpp does): init 0 = a
// product of single-site states uniform on the Bloch sphere, 1 = a Gaussian
// vector normalised (Haar); Haar circuits fix no parity.  psi [2^n] complex128.
void fe_initial_state(int n, int init_state, std::uint64_t trajectory_seed, std::vector<std::complex<double>> &psi)
{
    std::mt19937_64 rng(fe_splitmix64(trajectory_seed ^ 0x494e495449414c53ull));
    const std::size_t dim = std::size_t{1} << n;
    psi.assign(dim, {0.0, 0.0});
    if (init_state == 0)
    {
        std::uniform_real_distribution<double> unit(0.0, 1.0);
        psi[0] = 1.0;
        std::size_t active = 1;
        const double two_pi = 6.283185307179586476925286766559;
        for (int q = 0; q < n; ++q)
        {
            const double z = 1.0 - 2.0 * unit(rng);
            const double theta = std::acos(std::clamp(z, -1.0, 1.0));
            const double phi = two_pi * unit(rng);
            const std::complex<double> a0(std::cos(0.5 * theta), 0.0), a1 = std::polar(std::sin(0.5 * theta), phi);
            for (std::size_t i = 0; i < active; ++i)
            {
                const std::complex<double> prior = psi[i];
                psi[i] = prior * a0;
                psi[i | active] = prior * a1;
            }
            active <<= 1;
        }
        return;
    }
    std::normal_distribution<double> normal(0.0, 1.0);
    long double norm2 = 0.0L;
    for (std::size_t i = 0; i < dim; ++i)
    {
        const double re = normal(rng);
        const double im = normal(rng);
        psi[i] = {re, im};
        norm2 += static_cast<long double>(re * re + im * im);
    }
    const double inv = static_cast<double>(1.0L / std::sqrt(norm2));
    for (auto &x : psi) x *= inv;
}

bool file_exists(const std::string &path)
{
    struct stat sb;
    return ::stat(path.c_str(), &sb) == 0;
}

bool fe_records_entropy(int t, bool enabled, int max_t, int stride)
{
    if (!enabled || max_t < 0) return false;
    if (t == 0) return true;
    if (t > max_t) return false;
    return t % stride == 0 || t == max_t;
}

void fe_write_stat(std::ostream &out, const Welford &w)
{
    if (w.count == 0) out << "nan,nan,nan";
    else out << w.mean << ',' << w.sd() << ',' << w.se();
}

void fe_usage(const char *prog)
{
    std::cout << "Usage: " << prog << " [N] [p] [realizations] [circ_type] [init_state]\n\n"
              << "  N            sites of the periodic ring (even, <= 26)\n"
              << "  p            measurement rate in [0,1]\n"
              << "  realizations trajectories\n"
              << "  circ_type    1 = random Haar U(4) brickwork (the only one built here)\n"
              << "  init_state   0 = random product state, 1 = random Haar state\n\n"
              << "4L equilibration + 24L production timesteps.  Environment: MIPT_FREE_ENERGY_SEED,\n"
              << "MIPT_FREE_ENERGY_HALF_ENTROPY=1, MIPT_FREE_ENERGY_ENTROPY_MAX_T (default 2L),\n"
              << "MIPT_FREE_ENERGY_ENTROPY_STRIDE=1, MIPT_FREE_ENERGY_RESUME (0: overwrite existing CSVs;\n"
              << "resuming is not built here, so existing CSVs are otherwise refused).\n";
}

int run_free_energy(int argc, char **argv)
{
    if (argc > 1 && (std::strcmp(argv[1], "--help") == 0 || std::strcmp(argv[1], "-h") == 0))
    {
        fe_usage(argv[0]);
        return 0;
    }
    if (argc != 6)
    {
        fe_usage(argv[0]);
        return 2;
    }
    const int n = std::stoi(argv[1]);
    const double p = std::stod(argv[2]);
    const long parsed = std::stol(argv[3]);
    if (parsed < 1 || parsed > 1000000000L) throw std::invalid_argument("realizations must lie in [1,1000000000].");
    const int realizations = static_cast<int>(parsed);
    const int circ = std::stoi(argv[4]);
    const int init_state = std::stoi(argv[5]);
    if (circ != 1)
        throw std::invalid_argument("circ_type " + std::to_string(circ) +
                                    " is not built on this engine; use 1 (random Haar U(4)).");
    if (n < 2 || n >= 31) throw std::invalid_argument("N must lie in [2,30] for the dense exact-state protocol.");
    if (n % 2 != 0) throw std::invalid_argument("Random Haar Unitary requires an even N for periodic brickwork geometry.");
    if (n > 26) throw std::invalid_argument("this engine holds states up to N = 26.");
    if (!(p >= 0.0 && p <= 1.0)) throw std::invalid_argument("p must lie in [0,1].");
    if (init_state != 0 && init_state != 1) throw std::invalid_argument("init_state must be 0 or 1.");
    const int eq = 4 * n, prod = 24 * n, total = eq + prod;
    const bool record = env_flag("MIPT_FREE_ENERGY_HALF_ENTROPY", true);
    const int max_t = record ? env_int("MIPT_FREE_ENERGY_ENTROPY_MAX_T", 2 * n, 0, total) : -1;
    const int stride = env_int("MIPT_FREE_ENERGY_ENTROPY_STRIDE", 1, 1, total);
    std::uint64_t master;
    if (const char *e = std::getenv("MIPT_FREE_ENERGY_SEED"); e && *e)
    {
        char *end = nullptr;
        const long long v = std::strtoll(e, &end, 10);
        if (*end != '\0' || v < -1) throw std::invalid_argument("MIPT_FREE_ENERGY_SEED must be an integer >= -1.");
        master = v >= 0 ? static_cast<std::uint64_t>(v)
                        : (static_cast<std::uint64_t>(std::random_device{}()) << 32) ^ std::random_device{}();
    }
    else
        master = (static_cast<std::uint64_t>(std::random_device{}()) << 32) ^ std::random_device{}();
    const std::string prefix = "free_energy_haar_n_" + std::to_string(n) + "_p_" + number_for_filename(p) + "_init_" +
                               std::to_string(init_state) + "_reals_" +
                               compact_count(static_cast<std::uint64_t>(realizations));
    const std::string samples_path = prefix + "_samples.csv", series_path = prefix + "_timeseries.csv";
    if (env_flag("MIPT_FREE_ENERGY_RESUME", true) && (file_exists(samples_path) || file_exists(series_path)))
        throw std::runtime_error(samples_path + " or " + series_path +
                                 " exists; resuming is not built on this engine. Move it aside, or set "
                                 "MIPT_FREE_ENERGY_RESUME=0 to start over.");
    std::ofstream samples(samples_path, std::ios::out | std::ios::trunc);
    if (!samples) throw std::runtime_error("Could not open sample CSV: " + samples_path);
    samples << std::setprecision(17);
    samples << "N,p,circ_type,circuit_name,init_state,trajectory,parity_sector,trajectory_seed,equil_timesteps,"
               "production_timesteps,total_timesteps,measurements_total,measurements_production,F_total,"
               "F_production,production_rate_endpoint,production_slope,free_energy_density_tilde\n";
    std::vector<Welford> cum(total + 1), rate(total + 1), shalf(total + 1);
    std::vector<int> rec_t;
    for (int t = 1; t <= total; ++t)
        if (fe_records_entropy(t, record, max_t, stride)) rec_t.push_back(t);

    double budget_gib = 64.0;
    if (const char *e = std::getenv("MIPT_HIP_STATE_GIB")) budget_gib = std::max(0.001, std::atof(e));
    const int per_batch = static_cast<int>(
        std::max<double>(1.0, std::min<double>(realizations, budget_gib * 1073741824.0 / std::ldexp(8.0, n))));
    std::vector<std::complex<double>> psi;
    for (int first = 0; first < realizations; first += per_batch)
    {
        const int count = std::min(per_batch, realizations - first);
        std::vector<mipt_hip_traj_spec> specs(count);
        std::vector<std::uint64_t> tseed(count);
        for (int i = 0; i < count; ++i)
        {
            tseed[i] = fe_splitmix64(master ^ static_cast<std::uint64_t>(first + i));
            specs[i] = mipt_hip_traj_spec{};
            specs[i].L = n;
            specs[i].layers = total;
            specs[i].p = p;
            specs[i].seed = tseed[i];
            specs[i].closed = 1;
            specs[i].gate_family = 1;
            specs[i].probe_site = -1;
            specs[i].shared_gates = -1;
        }
        mipt_hip_batch *b = nullptr;
        const mipt_hip_batch_config cfg{0, 0, 0, 0};
        check(mipt_hip_batch_create(&cfg, nullptr, &b), "batch_create");
        std::vector<double> sh(count * (rec_t.size() + 1), 0.0);
        int st = mipt_hip_batch_setup(b, specs.data(), count);
        if (!st) st = mipt_hip_batch_generate(b);
        std::vector<float> ri;
        for (int i = 0; i < count && st == 0; ++i)
        {
            fe_initial_state(n, init_state, tseed[i], psi);
            ri.resize(2 * psi.size());
            for (std::size_t k = 0; k < psi.size(); ++k)
            {
                ri[2 * k] = static_cast<float>(psi[k].real());
                ri[2 * k + 1] = static_cast<float>(psi[k].imag());
            }
            st = mipt_hip_batch_set_state(b, i, ri.data());
        }
        if (!st && record) st = mipt_hip_batch_observables_all(b, 0, 0, sh.data(), nullptr);
        int cur = 0;
        for (std::size_t r = 0; r < rec_t.size() && st == 0; ++r)
        {
            st = mipt_hip_batch_evolve_range(b, cur, rec_t[r]);
            cur = rec_t[r];
            if (!st) st = mipt_hip_batch_synchronize(b);
            if (!st) st = mipt_hip_batch_observables_all(b, 0, 0, sh.data() + (r + 1) * count, nullptr);
        }
        if (!st && cur < total) st = mipt_hip_batch_evolve_range(b, cur, total);
        if (!st) st = mipt_hip_batch_synchronize(b);
        std::vector<int> out(static_cast<std::size_t>(total) * n);
        std::vector<std::uint8_t> flg(out.size());
        std::vector<double> p1(out.size());
        for (int i = 0; i < count && st == 0; ++i)
        {
            st = mipt_hip_batch_outcomes(b, i, out.data(), flg.data(), p1.data(), nullptr);
            if (st) break;
            double f = 0.0, fprod = 0.0;
            std::uint64_t meas = 0, meas_prod = 0;
            double sx = 0, sy = 0, sxx = 0, sxy = 0;
            std::uint64_t nfit = 0;
            cum[0].add(0.0);
            if (record) shalf[0].add(sh[i]);
            std::size_t r = 0;
            for (int t = 1; t <= total; ++t)
            {
                double layer = 0.0;
                int m = 0;
                for (int s2 = 0; s2 < n; ++s2)
                {
                    const std::size_t k = static_cast<std::size_t>(t - 1) * n + s2;
                    if (out[k] < 0) continue;
                    ++m;
                    layer -= std::log(out[k] ? p1[k] : 1.0 - p1[k]);
                }
                f += layer;
                meas += m;
                cum[t].add(f);
                rate[t].add(f / (static_cast<double>(t) * n));
                if (r < rec_t.size() && rec_t[r] == t)
                {
                    shalf[t].add(sh[(r + 1) * count + i]);
                    ++r;
                }
                if (t > eq)
                {
                    const double x = t - eq;
                    fprod += layer;
                    meas_prod += m;
                    ++nfit;
                    sx += x, sy += fprod, sxx += x * x, sxy += x * fprod;
                }
            }
            const double den = static_cast<double>(nfit) * sxx - sx * sx;
            if (nfit < 2 || !(den > 0.0)) throw std::runtime_error("Not enough production points for a free-energy slope.");
            const double slope = (static_cast<double>(nfit) * sxy - sx * sy) / den;
            samples << n << ',' << p << ',' << circ << ",\"Random Haar Unitary\"," << init_state << ',' << first + i
                    << ",-1," << tseed[i] << ',' << eq << ',' << prod << ',' << total << ',' << meas << ','
                    << meas_prod << ',' << f << ',' << fprod << ',' << fprod / prod << ',' << slope << ','
                    << slope / n << '\n';
        }
        (void)mipt_hip_batch_destroy(b);
        check(st, "free-energy batch");
        samples.flush();
        std::cerr << "realizations: " << first + count << '/' << realizations << '\n';
    }
    if (!samples) throw std::runtime_error("Failed while writing sample CSV: " + samples_path);
    std::ofstream series(series_path, std::ios::out | std::ios::trunc);
    if (!series) throw std::runtime_error("Could not open time-series CSV: " + series_path);
    series << std::setprecision(17);
    series << "N,p,circ_type,circuit_name,init_state,requested_realizations,completed_realizations,t,t_over_L,phase,"
              "F_mean,F_stddev,F_stderr,F_over_tL_mean,F_over_tL_stddev,F_over_tL_stderr,S_half_mean,S_half_stddev,"
              "S_half_stderr,S_half_count\n";
    for (int t = 0; t <= total; ++t)
    {
        const char *phase = t == 0 ? "initial" : (t <= eq ? "equilibration" : "production");
        series << n << ',' << p << ',' << circ << ",\"Random Haar Unitary\"," << init_state << ',' << realizations
               << ',' << cum[0].count << ',' << t << ',' << static_cast<double>(t) / n << ',' << phase << ',';
        fe_write_stat(series, cum[t]);
        series << ',';
        fe_write_stat(series, rate[t]);
        series << ',';
        fe_write_stat(series, shalf[t]);
        series << ',' << shalf[t].count << '\n';
    }
    series.flush();
    if (!series) throw std::runtime_error("Failed while writing time-series CSV: " + series_path);
    std::cout << "Saved " << samples_path << " and " << series_path << ".\n";
    return 0;
}

} // namespace

extern "C" {

// Per-result errors of trajectory i of batch b: its gates drawn again in FP64
// by the batch's counter-based sampler (the device rounds them to complex64),
// its flags and uniforms exported, the whole trajectory replayed on the plain
// engine in complex128, and the amplitudes, S(L/2) and TMI of that state (FP64
// Gram and eigenvalues) compared with the batch's.  A draw that comes out
// differently in the replay (only near its threshold) leaves the errors NaN.
static int fp64_replay_errors(mipt_hip_batch *b, mipt_hip_engine *eng128, int i, const mipt_hip_traj_spec &s,
                              const std::vector<int> &outcomes, int tmi_block, int tmi_cycles, const double *tmi_out,
                              mipt_hip_traj_result &r)
{
    using namespace mipt_hip;
    const int L = s.L, layers = s.layers, gx = s.grid_x, closed = s.closed ? 1 : 0;
    std::vector<double> g;
    for (int t = 0; t < layers; ++t)
        for (int k = 0; k < geo_bond_count(L, gx, t, closed); ++k)
        {
            double qr[16], qi[16];
            haar_u4_philox(s.seed, static_cast<std::uint32_t>(t), static_cast<std::uint32_t>(k), qr, qi);
            for (int e = 0; e < 16; ++e)
            {
                g.push_back(qr[e]);
                g.push_back(qi[e]);
            }
        }
    const std::size_t draws = static_cast<std::size_t>(layers) * L;
    std::vector<std::uint8_t> flags(draws);
    std::vector<double> unif(draws);
    int st = mipt_hip_batch_export_layers(b, i, nullptr, flags.data(), unif.data());
    if (st) return st;
    std::vector<int> out128(draws);
    if ((st = mipt_hip_engine_run_trajectory(eng128, L, layers, closed, gx, g.data(), flags.data(), unif.data(), 4, 0,
                                             10, out128.data())) != 0)
        return st;
    for (std::size_t k = 0; k < draws; ++k)
        if (out128[k] != outcomes[k]) return OK;  // diverged at a near-threshold draw: errors stay NaN
    const std::size_t dim = std::size_t{1} << L;
    std::vector<float> psi32(2 * dim);
    std::vector<double> psi64(2 * dim);
    if ((st = mipt_hip_batch_get_state(b, i, psi32.data())) != 0) return st;
    if ((st = mipt_hip_engine_to_host(eng128, psi64.data())) != 0) return st;
    double amp = 0.0;
    for (std::size_t k = 0; k < dim; ++k)
        amp = std::max(amp, std::hypot(psi32[2 * k] - psi64[2 * k], psi32[2 * k + 1] - psi64[2 * k + 1]));
    r.amp_err = amp;
    const void *dev = mipt_hip_engine_data(eng128);
    std::vector<int> kept(L / 2);
    for (int q = 0; q < L / 2; ++q) kept[q] = q;
    double s64 = 0.0;
    if ((st = mipt_hip_entropy_svd_f64(dev, L, kept.data(), L / 2, 0, &s64)) != 0) return st;
    r.s_half_err = std::abs(r.s_half - s64);
    if (tmi_cycles > 0)
    {
        double worst = 0.0;
        for (int c = 0; c < tmi_cycles; ++c)
        {
            double t64 = 0.0;
            if ((st = mipt_hip_tmi_f64(dev, L, L, tmi_block > 0 ? tmi_block : L / 4, c, &t64, nullptr)) != 0) return st;
            worst = std::max(worst, std::abs(tmi_out[static_cast<std::size_t>(i) * tmi_cycles + c] - t64));
        }
        r.tmi_err = worst;
    }
    return OK;
}

int mipt_hip_run_batch(const mipt_hip_traj_spec *specs, int count, int tmi_block, int tmi_cycles, int entropy_order,
                       double *tmi_out, mipt_hip_traj_result *results, void *hip_stream)
{
    if (!specs || count <= 0 || !results || tmi_cycles < 0 || (tmi_cycles > 0 && !tmi_out) ||
        (entropy_order != 1 && entropy_order != 2))
    {
        mipt_hip::set_error("run_batch: bad arguments");
        return mipt_hip::E_ARG;
    }
    mipt_hip_batch *b = nullptr;
    mipt_hip_engine *eng = nullptr, *eng128 = nullptr;
    int st = 0;
    const mipt_hip_batch_config cfg{0, 0, 0, 0};
    if ((st = mipt_hip_batch_create(&cfg, hip_stream, &b)) != 0) return st;
    do
    {
        if ((st = mipt_hip_batch_setup(b, specs, count)) != 0) break;
        if ((st = mipt_hip_batch_generate(b)) != 0) break;
        if ((st = mipt_hip_batch_evolve(b)) != 0) break;
        if ((st = mipt_hip_batch_synchronize(b)) != 0) break;
        if ((st = mipt_hip_engine_create(0, mipt_hip_batch_stream(b), &eng)) != 0) break;
        // S(L/2) and (von Neumann) TMI of all trajectories with batched eigensolves
        std::vector<double> sh(count), bloch(static_cast<std::size_t>(count) * 3);
        if ((st = mipt_hip_batch_observables_all(b, tmi_block, entropy_order == 1 ? tmi_cycles : 0, sh.data(),
                                                 entropy_order == 1 ? tmi_out : nullptr)) != 0)
            break;
        if ((st = mipt_hip_batch_probe_bloch(b, bloch.data())) != 0) break;
        for (int i = 0; i < count && st == 0; ++i)
        {
            const mipt_hip_traj_spec &s = specs[i];
            mipt_hip_traj_result &r = results[i];
            const std::size_t draws = static_cast<std::size_t>(s.layers) * s.L;
            std::vector<int> out(draws);
            std::vector<std::uint8_t> flg(draws);
            if ((st = mipt_hip_batch_outcomes(b, i, out.data(), flg.data(), nullptr, nullptr)) != 0) break;
            r.measured = 0;
            r.flagged_draws = 0;
            for (std::size_t k = 0; k < draws; ++k)
            {
                r.measured += out[k] >= 0;
                r.flagged_draws += flg[k] != 0;
            }
            r.s_half = sh[i];
            for (int k = 0; k < 3; ++k) r.bloch[k] = bloch[static_cast<std::size_t>(i) * 3 + k];
            if (s.probe_site < 0)
                r.probe_entropy = NAN;
            else
            {
                // cpp/include/mipt/ancilla_mi.hpp:24-52 (nats)
                const double len = std::min(1.0, std::sqrt(r.bloch[0] * r.bloch[0] + r.bloch[1] * r.bloch[1] +
                                                           r.bloch[2] * r.bloch[2]));
                const double l0 = 0.5 * (1 + len), l1 = 0.5 * (1 - len);
                r.probe_entropy = (l0 > 0 ? -l0 * std::log(l0) : 0.0) + (l1 > 0 ? -l1 * std::log(l1) : 0.0);
            }
            if (entropy_order == 2)
                for (int c = 0; c < tmi_cycles && st == 0; ++c)
                    st = mipt_hip_batch_tmi(b, i, tmi_block, c, 2, &tmi_out[static_cast<std::size_t>(i) * tmi_cycles + c]);
            if (st != 0) break;
            const int nq = s.L + (s.probe_site >= 0 ? 1 : 0);
            double sq = 0.0;
            if ((st = mipt_hip_engine_adopt(eng, mipt_hip_batch_state_ptr(b, i), nq)) != 0) break;
            if ((st = mipt_hip_engine_squared_norm(eng, &sq)) != 0) break;
            r.norm_error = std::abs(sq - 1.0);
            r.amp_err = r.s_half_err = r.tmi_err = NAN;
            const bool replayable = s.probe_site < 0 && s.probe_count == 0 && s.p_zero_layers == 0 && s.shared_gates < 0;
            if (s.fp64_check && replayable)
            {
                if (!eng128 && (st = mipt_hip_engine_create(1, mipt_hip_batch_stream(b), &eng128)) != 0) break;
                if ((st = fp64_replay_errors(b, eng128, i, s, out, tmi_block, entropy_order == 1 ? tmi_cycles : 0,
                                             tmi_out, r)) != 0)
                    break;
            }
        }
    } while (false);
    if (eng) (void)mipt_hip_engine_destroy(eng);
    if (eng128) (void)mipt_hip_engine_destroy(eng128);
    (void)mipt_hip_batch_destroy(b);
    return st;
}

int mipt_hip_entropy_main(int argc, char **argv)
{
    try
    {
        return run_entropy(argc, argv);
    }
    catch (const std::exception &e)
    {
        std::cerr << "\nerror: " << e.what() << '\n';
        return 1;
    }
}

int mipt_hip_mipt_main(int argc, char **argv)
{
    try
    {
        return run_mipt(argc, argv);
    }
    catch (const std::exception &e)
    {
        std::cerr << "\nerror: " << e.what() << '\n';
        return 1;
    }
}

int mipt_hip_free_energy_initial_state(int n, int init_state, uint64_t trajectory_seed, double *psi_ri)
{
    if (n < 1 || n > 26 || (init_state != 0 && init_state != 1) || !psi_ri)
    {
        mipt_hip::set_error("free_energy_initial_state: bad arguments");
        return mipt_hip::E_ARG;
    }
    std::vector<std::complex<double>> psi;
    fe_initial_state(n, init_state, trajectory_seed, psi);
    for (std::size_t k = 0; k < psi.size(); ++k)
    {
        psi_ri[2 * k] = psi[k].real();
        psi_ri[2 * k + 1] = psi[k].imag();
    }
    return 0;
}

int mipt_hip_free_energy_main(int argc, char **argv)
{
    try
    {
        return run_free_energy(argc, argv);
    }
    catch (const std::exception &e)
    {
        std::cerr << "free_energy: " << e.what() << '\n';
        return 1;
    }
}

int mipt_hip_probed_main(int argc, char **argv)
{
    try
    {
        return run_probed(argc, argv);
    }
    catch (const std::exception &e)
    {
        std::cerr << "\nerror: " << e.what() << '\n';
        return 1;
    }
}

int mipt_hip_sim_tmi_main(int argc, char **argv)
{
    try
    {
        return run_sim_tmi(argc, argv);
    }
    catch (const std::exception &e)
    {
        std::cerr << "\nerror: " << e.what() << '\n';
        return 1;
    }
}

} // extern "C"
