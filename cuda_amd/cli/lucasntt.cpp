// lucasntt -- command-line Lucas-Lehmer / PRP tester on the MI355X engine (liblnt).
//
// Drop-in for the reference program's surface: the LucasNTT.ini keys, the positional
// arguments, the .dat/.txt/.env files, resume after an interruption, the console
// messages and the verdict (main.cpp:22-39, main_process.cu:64-196, lucas_prp.cu:137-346).
// The hot path is liblnt's C ABI (include/lnt_api.h).
//
//   lucasntt [exponent [iterations [output_base [fft]]]]
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cinttypes>
#include <csignal>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "../../include/lnt_api.h"
#include "ini.hpp"

// ------------------------------------------------------------------ console logger
// Verbosity levels of the reference Logger (common/common.cu:19-112).
enum class Verbosity { Verbose, Normal, Silent };
static Verbosity g_verbosity = Verbosity::Normal;

static void WriteLine(Verbosity level, const char *fmt, ...) {
    if (level < g_verbosity) return;
    va_list ap;
    va_start(ap, fmt);
    std::vprintf(fmt, ap);
    va_end(ap);
    std::printf("\n");
    std::fflush(stdout);
}
#define SAY(...) WriteLine(Verbosity::Normal, __VA_ARGS__)
#define VERBOSE(...) WriteLine(Verbosity::Verbose, __VA_ARGS__)

static std::atomic<bool> g_interrupted{false};
static void CtrlHandler(int) { g_interrupted = true; }  // main_process.cu:43-46

// Y/N prompt for a length that is not the secured one (main_process.cu:49-62).  End of
// input counts as "N" so the program cannot spin on a closed stdin.
static bool Confirm(const std::string &question) {
    std::string answer;
    while (true) {
        std::cout << question << " (Y/N) : " << std::flush;
        if (!std::getline(std::cin, answer)) return false;
        if (!answer.empty()) {
            char c = (char)std::toupper((unsigned char)answer[0]);
            if (c == 'Y') return true;
            if (c == 'N') return false;
        }
        std::cout << "Please enter Y or N.\n";
    }
}

// NttLengths::StringToLength / FillLength / StringToLengths (ntt_lengths.cu:32-75).
struct Length {
    uint32_t n = 1, log2_n = 0, log3_n = 0, log5_n = 0;
};
static Length StringToLength(std::string d) {
    uint32_t factor = 1;
    if (!d.empty() && std::toupper((unsigned char)d.back()) == 'K') {
        factor = 1024;
        d.pop_back();
    }
    if (!d.empty() && std::toupper((unsigned char)d.back()) == 'M') {
        factor = 1024 * 1024;
        d.pop_back();
    }
    Length l;
    l.n = (uint32_t)std::stoi(d) * factor;
    uint32_t n = l.n;
    while (n && n % 2 == 0) l.log2_n++, n /= 2;
    while (n && n % 3 == 0) l.log3_n++, n /= 3;
    while (n && n % 5 == 0) l.log5_n++, n /= 5;
    if (n != 1) throw std::runtime_error("Invalid NTT length provided! Should a multiple of 2, 3 and 5.\n\n");
    if (l.log3_n > 1) throw std::runtime_error("Invalid NTT length provided! Cannot use exponent of 3 greater than 1.\n\n");
    if (l.log5_n > 1) throw std::runtime_error("Invalid NTT length provided! Cannot use exponent of 5 greater than 1.\n\n");
    return l;
}
static std::vector<Length> StringToLengths(const std::string &d) {
    std::vector<Length> out;
    size_t pos = 0;
    while (true) {
        size_t q = d.find(':', pos);
        out.push_back(StringToLength(d.substr(pos, q == std::string::npos ? std::string::npos : q - pos)));
        if (q == std::string::npos) break;
        pos = q + 1;
    }
    return out;
}
static std::string LengthsToString(const std::vector<Length> &ls) {
    std::string r;
    for (auto &l : ls) r += (r.empty() ? "" : ":") + std::to_string(l.n);
    return r;
}

// ------------------------------------------------------------------ configuration
struct Config {
    bool PRP = false;
    uint32_t exponent = 9689;
    uint32_t iterations = 0;
    uint32_t fromIteration = 1;
    bool wasInterrupted = false;
    bool dumpResidue = false, dumpStats = true, dumpEnvironment = false;
    std::string dumpFilename, statFilename, envFilename;
    std::string preferredPool, preferredKernels;
    std::string fft;  // provided length, "" = chosen by the length rule
    int gpu = 0;
    long performanceLoops = 500;
    long dumpIterations = 0;   // residue dump every this many iterations (0: about hourly)
    std::string confirmLength; // answer to the length question: "Y", "N" or "" (ask on stdin)
};

// MainProcess::MainProcess (main_process.cu:64-149).
static Config ReadConfig(int argc, char **argv) {
    Config c;
    Ini ini;
    bool have_ini = ini.LoadFile("LucasNTT.ini");
    std::string verbosity = have_ini ? ini.GetValue("Verbosity", "Normal") : "Normal";
    if (verbosity == "Verbose") g_verbosity = Verbosity::Verbose;
    if (verbosity == "Silent") g_verbosity = Verbosity::Silent;
    SAY("Welcome to LucasNTT, a pure-integer Lucas-Lehmer and PRP primality test for Mersenne numbers running on GPU.");
    SAY("------------------------------------------------------------------------------------------------------------");
    SAY("");
    c.gpu = have_ini ? (int)ini.GetLongValue("GPU", 0) : 0;
    c.performanceLoops = have_ini ? ini.GetLongValue("PerformanceLoops", 500) : 500;
    c.dumpFilename = argc > 3 ? std::string(argv[3]) + ".dat" : have_ini ? ini.GetValue("Dump", "LucasNTT_results.dat") : "LucasNTT_results.dat";
    c.statFilename = argc > 3 ? std::string(argv[3]) + ".txt" : have_ini ? ini.GetValue("Stats", "LucasNTT_results.txt") : "LucasNTT_results.txt";
    c.envFilename = argc > 3 ? std::string(argv[3]) + ".env" : have_ini ? ini.GetValue("Env", "LucasNTT_results.env") : "LucasNTT_results.env";
    VERBOSE("Filename for final results: %s", c.statFilename.c_str());
    c.dumpResidue = have_ini && ini.GetLongValue("dumpResidue", 0) == 1;
    c.dumpStats = have_ini ? ini.GetLongValue("dumpStats", 1) == 1 : true;
    c.dumpEnvironment = have_ini && ini.GetLongValue("dumpEnvironment", 0) == 1;
    // keys of this program only: the interval of the safety dump, and a non-interactive
    // answer to the question asked when the provided length is not the secured one
    c.dumpIterations = have_ini ? ini.GetLongValue("DumpIterations", 0) : 0;
    if (have_ini) {
        std::string a = ini.GetValue("ConfirmLength", "");
        if (!a.empty()) {
            char ch = (char)std::toupper((unsigned char)a[0]);
            if (ch != 'Y' && ch != 'N') throw std::runtime_error("ConfirmLength must be Y or N");
            c.confirmLength = std::string(1, ch);
        }
    }

    Ini last;
    c.wasInterrupted = last.LoadFile(c.statFilename) && last.GetLongValue("Interrupted", 0) == 1;
    const Ini *src = have_ini ? &ini : nullptr;
    if (c.wasInterrupted) {
        src = &last;  // the stats file replaces LucasNTT.ini (main_process.cu:101-106)
        SAY("Last execution was interrupted, trying to restore last execution status...");
    }
    c.PRP = src && src->GetLongValue("PRP", 0) == 1;
    c.exponent = argc > 1 ? (uint32_t)std::stoi(argv[1]) : src ? (uint32_t)src->GetLongValue("Exponent", 9689) : 9689;
    SAY("Exponent = %d", c.exponent);
    uint32_t maxIterations = c.PRP ? c.exponent : c.exponent - 2;
    c.iterations = argc > 2 ? (uint32_t)std::stoi(argv[2]) : src ? (uint32_t)src->GetLongValue("Iterations", maxIterations) : maxIterations;
    if (c.iterations > maxIterations || c.iterations <= 0) c.iterations = maxIterations;
    c.fromIteration = c.wasInterrupted && src ? (uint32_t)src->GetLongValue("currentIteration", 1) : 1;
    c.preferredPool = src ? src->GetValue("Pool", "") : "";
    c.preferredKernels = ";" + (src ? src->GetValue("Kernels", "") : std::string()) + ";";
    c.fft = argc > 4 ? std::string(argv[4]) : src ? src->GetValue("fft", "") : "";
    return c;
}

// ------------------------------------------------------------------ the test driver
class LucasPRP {
   public:
    LucasPRP(Config &cfg) : cfg_(cfg) {}
    ~LucasPRP() {
        if (ctx_) lnt_destroy(ctx_);
    }

    void Initialize() {
        std::vector<uint32_t> factors;
        if (!cfg_.fft.empty()) {
            lengths_ = StringToLengths(cfg_.fft);
            for (auto &l : lengths_) factors.push_back(l.n);
        } else {
            uint32_t l = lnt_secured_log2_length(cfg_.exponent);
            Length len;
            len.n = 1u << l;
            len.log2_n = l;
            lengths_.push_back(len);
            factors.push_back(len.n);
        }
        int rc = lnt_create(cfg_.gpu, cfg_.exponent, factors.data(), (int)factors.size(),
                            cfg_.PRP ? LNT_MODE_PRP : LNT_MODE_LL, &ctx_);
        if (rc != LNT_OK) throw std::runtime_error(std::string("lnt_create failed: ") + lnt_last_error(nullptr));
        n_ = lnt_length(ctx_);
        host_x_.assign(n_, 0);
        char buf[1024];
        lnt_describe(ctx_, buf, sizeof buf);
        std::string d = buf;
        size_t pos = 0;
        while (true) {
            size_t q = d.find(';', pos);
            names_.push_back(d.substr(pos, q == std::string::npos ? std::string::npos : q - pos));
            if (q == std::string::npos) break;
            pos = q + 1;
        }
        if (!cfg_.preferredPool.empty() && cfg_.preferredPool != names_[0])
            VERBOSE("Pool %s is a reference kernel chain; this engine runs %s", cfg_.preferredPool.c_str(), names_[0].c_str());
        if (cfg_.preferredKernels != ";;")
            VERBOSE("Kernels %s name reference kernels; this engine chooses its own", cfg_.preferredKernels.c_str());
    }

    // Autotune: replaces Process::GetBestPerformers (process.cu:45-127); every candidate is
    // validated against the others before its time counts.
    void Tune() {
        if (cfg_.performanceLoops <= 0) return;
        check(lnt_tune(ctx_, (int)cfg_.performanceLoops), "lnt_tune");
        SaveTuning();
        char buf[1024];
        lnt_describe(ctx_, buf, sizeof buf);
        names_.clear();
        std::string d = buf;
        size_t pos = 0;
        while (true) {
            size_t q = d.find(';', pos);
            names_.push_back(d.substr(pos, q == std::string::npos ? std::string::npos : q - pos));
            if (q == std::string::npos) break;
            pos = q + 1;
        }
        // every candidate's time, kept like the reference's performance map (name -> ms,
        // common/common.cu:83-112): console table and "Candidate Performance" stats lines
        std::vector<char> rep(1 << 16);
        check(lnt_tune_results(ctx_, rep.data(), rep.size()), "lnt_tune_results");
        std::vector<std::pair<std::string, double>> perf;
        std::string text = rep.data();
        size_t pos2 = 0;
        while (pos2 < text.size()) {
            size_t e = text.find('\n', pos2);
            std::string line = text.substr(pos2, e == std::string::npos ? std::string::npos : e - pos2);
            pos2 = e == std::string::npos ? text.size() : e + 1;
            size_t eq = line.rfind(" = ");
            if (eq == std::string::npos) continue;
            perf.push_back({line.substr(0, eq), std::atof(line.c_str() + eq + 3)});
        }
        std::sort(perf.begin(), perf.end());
        size_t width = 0;
        for (auto &pr : perf) width = std::max(width, pr.first.size());
        SAY("\nSummary of the performances of the different processes on %ld loops:", cfg_.performanceLoops);
        perfLines_.clear();
        for (auto &pr : perf) {
            SAY("%-*s = %8.3fms", (int)width, pr.first.c_str(), pr.second);
            perfLines_.push_back("Candidate Performance (in ms for " + std::to_string(cfg_.performanceLoops) +
                                 " loops): " + pr.first + " = " + fmt_g(pr.second));
        }
        // the chosen configuration's time per PerformanceLoops iterations paces the pit stops
        performance_ms_ = 0;
        for (auto &pr : perf)
            if (performance_ms_ == 0 || pr.second < performance_ms_) performance_ms_ = pr.second;
    }

    // The tuner's choices persist in LucasNTT.tune (one line per device and length), so a
    // later run with PerformanceLoops = 0 starts on the tuned plan.
    static constexpr const char *kTuneFile = "LucasNTT.tune";
    static void LoadTuning() {
        FILE *f = std::fopen(kTuneFile, "r");
        if (!f) return;
        std::string text;
        char buf[512];
        while (std::fgets(buf, sizeof buf, f)) text += buf;
        std::fclose(f);
        if (lnt_tune_import(text.c_str()) != LNT_OK)
            VERBOSE("Ignoring %s: %s", kTuneFile, lnt_last_error(nullptr));
    }
    static void SaveTuning() {
        std::vector<char> buf(1 << 16);
        if (lnt_tune_export(buf.data(), buf.size()) != LNT_OK) return;
        FILE *f = std::fopen(kTuneFile, "w");
        if (!f) return;
        std::fputs(buf.data(), f);
        std::fclose(f);
    }

    void Run() {
        SAY("\nThis test will run using the following processes:");
        SAY("MainProcess");
        SAY("    LucasPRP %s", LengthsToString(lengths_).c_str());
        for (size_t i = 0; i < names_.size(); i++) SAY("%s%s", i ? "            " : "        ", names_[i].c_str());
        SAY("\n\n----------------------------------------------");
        SAY("Now testing M(q) = 2^q - 1 where q = %d", cfg_.exponent);
        SAY(cfg_.PRP ? "With PRP pseudo-primality test" : "With Lucas-Lehmer primality test");
        SAY("From iteration %d to %d\n", cfg_.fromIteration, cfg_.iterations);
        std::signal(SIGINT, CtrlHandler);

        // lucas_prp.cu:137-149: resume or start from 4 / 3
        if (cfg_.wasInterrupted) {
            if (LoadResidue())
                SAY("Previous residue successfully loaded!");
            else {
                SAY("Failed to restore previous residue, restarting from iteration 1!");
                InitializeTest();
                cfg_.fromIteration = 1;
            }
        } else
            InitializeTest();

        auto t0 = std::chrono::steady_clock::now();
        check(lnt_sync(ctx_, nullptr), "lnt_sync");
        MainLoop();
        float ms = 0;
        check(lnt_sync(ctx_, &ms), "lnt_sync");
        totalTime_ = (float)std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();

        check(lnt_get_digits(ctx_, host_x_.data(), n_), "lnt_get_digits");
        uint32_t maxIterations = cfg_.PRP ? cfg_.exponent : cfg_.exponent - 2;
        if (currentIteration_ > maxIterations) {  // lucas_prp.cu:167-186
            isPrime_ = host_x_[0] == (uint64_t)(cfg_.PRP ? 9 : 0);
            for (uint32_t i = 1; i < n_ && isPrime_; i++)
                if (host_x_[i] != 0) isPrime_ = false;
            if (isPrime_)
                SAY("\n\nM(%d) is %sprime!\n", cfg_.exponent, cfg_.PRP ? "probably " : "");
            else {
                isComposite_ = true;
                SAY("\n\nM(%d) is composite!\n", cfg_.exponent);
            }
        } else
            SAY("\n\n%d loops performed with M(%d)\n", currentIteration_ - cfg_.fromIteration, cfg_.exponent);
        uint64_t r64 = 0;
        check(lnt_res64(ctx_, &r64), "lnt_res64");
        SAY("Res64 = %016" PRIX64, r64);
        res64_ = r64;
        if (cfg_.dumpResidue || g_interrupted || currentIteration_ <= cfg_.exponent - 2) DumpResidue();
        if (cfg_.dumpEnvironment) DumpEnvironment();
        if (cfg_.dumpStats) DumpStats();
    }

   private:
    static std::string fmt_g(double v) {
        char b[64];
        std::snprintf(b, sizeof b, "%g", v);
        return b;
    }
    void check(int rc, const char *what) {
        if (rc != LNT_OK) throw std::runtime_error(std::string(what) + " failed: " + lnt_last_error(ctx_));
    }

    void InitializeTest() {  // lucas_prp.cu:127-134
        std::fill(host_x_.begin(), host_x_.end(), 0);
        host_x_[0] = cfg_.PRP ? 3 : 4;
        check(lnt_set_digits(ctx_, host_x_.data(), n_), "lnt_set_digits");
    }

    // lucas_prp.cu:197-225.  Iterations are enqueued in batches that end at the next pit
    // stop, hourly dump or interruption check, so the order of events is the reference's.
    void MainLoop() {
        const double perIterMs = performance_ms_ > 0 ? performance_ms_ / cfg_.performanceLoops : 0.05;
        uint64_t pitStop = (uint64_t)(60.0 * 1000.0 / perIterMs * 0.98);
        uint64_t dumpIteration = (uint64_t)(3600.0 * 1000.0 / perIterMs * 0.98);
        if (cfg_.dumpIterations > 0) dumpIteration = (uint64_t)cfg_.dumpIterations;
        if (pitStop == 0) pitStop = 1;
        if (dumpIteration == 0) dumpIteration = 1;
        const uint64_t batch_cap = 4096;
        uint64_t j = 0, k = 0;
        start_ = std::chrono::steady_clock::now();
        currentIteration_ = cfg_.fromIteration;
        while (currentIteration_ <= cfg_.iterations) {
            if (g_interrupted) break;
            j++;
            if (j == dumpIteration) {
                j = 0;
                SAY("\nNow dumping intermediate residue for safety reasons...\n");
                check(lnt_get_digits(ctx_, host_x_.data(), n_), "lnt_get_digits");
                DumpResidue();
                DumpStats();
            }
            // this iteration plus as many following ones as fit before the next event.  The
            // dump happens before the iteration whose count j reaches dumpIteration, so the
            // following iterations of this batch may take j up to dumpIteration - 1 only
            // (j counts this iteration already and is below dumpIteration here).
            uint64_t left = (uint64_t)cfg_.iterations - currentIteration_ + 1;
            uint64_t until_dump = dumpIteration - 1 - j;
            uint64_t until_pit = pitStop - k - 1;
            uint64_t b = 1 + std::min<uint64_t>({left - 1, until_dump, until_pit, batch_cap - 1});
            check(lnt_iterate(ctx_, (uint32_t)b), "lnt_iterate");
            j += b - 1;
            k += b;
            currentIteration_ += (uint32_t)b;
            if (k == pitStop && g_verbosity != Verbosity::Silent) {
                k = 0;
                DisplayPitStop();
            }
        }
    }

    void DisplayPitStop() {  // lucas_prp.cu:231-252
        check(lnt_sync(ctx_, nullptr), "lnt_sync");
        double t = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - start_).count();
        long long h = (long long)(t / 3600000.0);
        long long m = (long long)((t / 1000.0 - h * 3600) / 60);
        long long s = (long long)(t / 1000.0) - h * 3600 - m * 60;
        uint32_t cur = currentIteration_ - 1;  // last iteration performed
        double done = (double)cur - cfg_.fromIteration + 1.0;
        double perIter = t / done;
        double remaining = perIter * ((double)cfg_.iterations - cur);
        long long h2 = (long long)(remaining / 3600000.0);
        long long m2 = (long long)((remaining / 1000.0 - h2 * 3600) / 60);
        long long s2 = (long long)(remaining / 1000.0) - h2 * 3600 - m2 * 60;
        double pct = 100.0 * cur / cfg_.iterations;
        SAY("%d%% - Elapsed %02lld:%02lld:%02lld - Remaining %02lld:%02lld:%02lld - Per iteration %.4f ms", (int)pct,
            h, m, s, h2, m2, s2, perIter);
    }

    // .dat: n lines "0x%016" PRIX64 "\n" in natural digit order (lucas_prp.cu:254-263).
    void DumpResidue() {
        FILE *fp = std::fopen(cfg_.dumpFilename.c_str(), "w+");
        if (!fp) throw std::runtime_error("cannot write " + cfg_.dumpFilename);
        for (uint32_t i = 0; i < n_; i++) std::fprintf(fp, "0x%016" PRIX64 "\n", host_x_[i]);
        std::fclose(fp);
    }

    // lucas_prp.cu:326-346: 19-byte records, then the file is removed.
    bool LoadResidue() {
        FILE *fp = std::fopen(cfg_.dumpFilename.c_str(), "r");
        if (!fp) return false;
        char buf[20];
        uint32_t i = 0;
        std::fill(host_x_.begin(), host_x_.end(), 0);
        while (std::fread(buf, 1, 19, fp) > 0) {
            buf[19] = 0;
            if (i < n_) host_x_[i] = (uint64_t)std::strtoull(buf, nullptr, 16);
            i++;
        }
        std::fclose(fp);
        std::remove(cfg_.dumpFilename.c_str());
        check(lnt_set_digits(ctx_, host_x_.data(), n_), "lnt_set_digits");
        return true;
    }

    void DumpEnvironment() {  // lucas_prp.cu:265-291, HIP device properties
        FILE *fp = std::fopen(cfg_.envFilename.c_str(), "w+");
        if (!fp) return;
        int rt = 0, devCount = 0;
        (void)hipRuntimeGetVersion(&rt);
        (void)hipGetDeviceCount(&devCount);
        std::fprintf(fp, "HIP Version = %d\n", rt);
        std::fprintf(fp, "HIP devices = %d\n", devCount);
        for (int i = 0; i < devCount; i++) {
            hipDeviceProp_t pr;
            if (hipGetDeviceProperties(&pr, i) != hipSuccess) continue;
            std::fprintf(fp, "Device %d = %s\n", i, pr.name);
            std::fprintf(fp, "  Architecture = %s\n", pr.gcnArchName);
            std::fprintf(fp, "  Multiprocessors = %d\n", pr.multiProcessorCount);
            std::fprintf(fp, "  Global memory (mb) = %d\n", (int)(pr.totalGlobalMem / (1024 * 1024)));
            std::fprintf(fp, "  Shared memory (kb) = %d\n", (int)(pr.sharedMemPerBlock / 1024));
            std::fprintf(fp, "  Constant memory (kb) = %d\n", (int)(pr.totalConstMem / 1024));
            std::fprintf(fp, "  Block registers = %d\n", pr.regsPerBlock);
            std::fprintf(fp, "  Warp size = %d\n", pr.warpSize);
            std::fprintf(fp, "  Threads per block = %d\n", pr.maxThreadsPerBlock);
            std::fprintf(fp, "  Max block dimensions = [%d, %d, %d]\n", pr.maxThreadsDim[0], pr.maxThreadsDim[1], pr.maxThreadsDim[2]);
            std::fprintf(fp, "  Max grid dimensions = [%d, %d, %d]\n", pr.maxGridSize[0], pr.maxGridSize[1], pr.maxGridSize[2]);
        }
        std::fclose(fp);
    }

    // Stats file, key order of lucas_prp.cu:293-324, plus Res64.
    void DumpStats() {
        FILE *fp = std::fopen(cfg_.statFilename.c_str(), "w+");
        if (!fp) throw std::runtime_error("cannot write " + cfg_.statFilename);
        std::fprintf(fp, "Exponent = %d\n", cfg_.exponent);
        if (isPrime_) std::fprintf(fp, "Conclusion = %sPrime\n", cfg_.PRP ? "Probably " : "");
        if (isComposite_) std::fprintf(fp, "Conclusion = Composite\n");
        if (!isPrime_ && !isComposite_) std::fprintf(fp, "Conclusion = Partial\n");
        std::fprintf(fp, "PRP = %d\n", cfg_.PRP ? 1 : 0);
        std::fprintf(fp, "Iterations = %d\n", cfg_.iterations);
        std::fprintf(fp, "fft = %s\n", LengthsToString(lengths_).c_str());
        uint32_t l2 = 0, l3 = 0, l5 = 0;
        for (auto &l : lengths_) l2 += l.log2_n, l3 += l.log3_n, l5 += l.log5_n;
        std::fprintf(fp, "n = %d\n", n_);
        std::fprintf(fp, "log2_n = %d\n", l2);
        std::fprintf(fp, "log3_n = %d\n", l3);
        std::fprintf(fp, "log5_n = %d\n", l5);
        for (auto &l : perfLines_) std::fprintf(fp, "%s\n", l.c_str());
        std::fprintf(fp, "Pool = %s\n", names_[0].c_str());
        for (size_t i = 1; i < names_.size(); i++) std::fprintf(fp, "Kernel%d = %s\n", (int)(i - 1), names_[i].c_str());
        std::fprintf(fp, "Interrupted = %d\n", g_interrupted || currentIteration_ < cfg_.iterations ? 1 : 0);
        std::fprintf(fp, "currentIteration = %d\n", currentIteration_);
        uint32_t done = currentIteration_ - cfg_.fromIteration + 1;
        std::fprintf(fp, "Per iteration (ms) = %g\n", totalTime_ / (done ? done : 1));
        if (cfg_.wasInterrupted)
            std::fprintf(fp, "Warning! total time is not relevant since the full process has been interrupted by CTRL-C\n");
        std::fprintf(fp, "Total time (sec) = %g\n", totalTime_ / 1000);
        std::fprintf(fp, "Res64 = %016" PRIX64 "\n", res64_);
        std::fclose(fp);
    }

    Config &cfg_;
    lnt_ctx *ctx_ = nullptr;
    uint32_t n_ = 0;
    std::vector<uint64_t> host_x_;
    std::vector<Length> lengths_;
    std::vector<std::string> names_;
    std::vector<std::string> perfLines_;
    double performance_ms_ = 0;
    float totalTime_ = 0;
    uint32_t currentIteration_ = 1;
    bool isPrime_ = false, isComposite_ = false;
    uint64_t res64_ = 0;
    std::chrono::steady_clock::time_point start_;
};

int main(int argc, char **argv) {  // main.cpp:22-39
    try {
        Config cfg = ReadConfig(argc, argv);
        if (!cfg.fft.empty()) {
            uint32_t log_n = 0;
            for (auto &l : StringToLengths(cfg.fft)) log_n += l.log2_n;
            int diff = (int)lnt_secured_log2_length(cfg.exponent) - (int)log_n;  // CheckProvidedLength
            if (diff != 0) {
                const char *q = diff > 0 ? "\n\nThe provided length is too small. Due to the risk of overflow, the result will not be reliable.\nDo you still want to process?"
                                         : "\n\nThe provided length is not optimal. The process can be accelerated.\nDo you still want to process?";
                bool yes;
                if (cfg.confirmLength.empty()) yes = Confirm(q);
                else {  // answered by the ConfirmLength key: no prompt, stdin is not read
                    yes = cfg.confirmLength == "Y";
                    SAY("%s (Y/N) : %s (ConfirmLength)", q, cfg.confirmLength.c_str());
                }
                if (!yes) throw std::runtime_error("Process cancelled by user. Invalid FFT length.");
            }
        }
        if (std::getenv("LNT_DRY_RUN")) {  // print the resolved configuration only
            std::printf("DRYRUN exponent=%u PRP=%d iterations=%u fromIteration=%u wasInterrupted=%d fft=%s "
                        "dump=%s stats=%s env=%s dumpResidue=%d dumpStats=%d dumpEnvironment=%d gpu=%d loops=%ld "
                        "pool=%s kernels=%s dumpIterations=%ld confirmLength=%s\n",
                        cfg.exponent, cfg.PRP, cfg.iterations, cfg.fromIteration, cfg.wasInterrupted, cfg.fft.c_str(),
                        cfg.dumpFilename.c_str(), cfg.statFilename.c_str(), cfg.envFilename.c_str(), cfg.dumpResidue,
                        cfg.dumpStats, cfg.dumpEnvironment, cfg.gpu, cfg.performanceLoops, cfg.preferredPool.c_str(),
                        cfg.preferredKernels.c_str(), cfg.dumpIterations,
                        cfg.confirmLength.empty() ? "ask" : cfg.confirmLength.c_str());
            return 0;
        }
        SAY("");
        LucasPRP::LoadTuning();
        LucasPRP test(cfg);
        test.Initialize();
        test.Tune();
        test.Run();
        VERBOSE("End.");
    } catch (const std::runtime_error &e) {
        std::printf("%s\n", e.what());
        return 1;
    } catch (...) {
        std::printf("Unhandled exception! Abort!\n");
        return 1;
    }
    return 0;
}
