// Regression checks of the reference (tests/sim_tests.cpp of
// chdnl0420-svg/stardust) restated against the C++ `Sim` shim over this
// library. The reference seeds its scenes with presets that its current
// header no longer has (SpiralDisk); here the same scenes are set up through
// Sim::setParticles with a deterministic disk / cloud generator. Thresholds
// are the reference's. Exit 2 when no device is visible, as the reference.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../stardust_amd/cpp/Sim.h"

static int g_pass = 0, g_fail = 0;

static void check(bool ok, const std::string& name, const std::string& detail) {
    if (ok) {
        ++g_pass;
        std::printf("  [PASS] %-52s %s\n", name.c_str(), detail.c_str());
    } else {
        ++g_fail;
        std::printf("  [FAIL] %-52s %s\n", name.c_str(), detail.c_str());
    }
    std::fflush(stdout);
}

// SplitMix64 -> uniform doubles in [0, 1).
struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    double next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (double)(z >> 11) * (1.0 / 9007199254740992.0);
    }
};

// Exponential disk of radius ~0.2 box centred at 0.5 with circular speeds for
// the given reference gravity (G_N,box * M = gravity / G on the isolated mesh).
static void spiral_disk(int n, int G, float gravity, uint64_t seed, std::vector<float>& p, std::vector<float>& v) {
    p.assign((size_t)n * 4, 0.f);
    v.assign((size_t)n * 4, 0.f);
    Rng r(seed);
    const double Rd = 0.05, Rmax = 0.22, h = 0.004;
    const double GM = gravity / (double)G;
    const double cmax = 1.0 - (1.0 + Rmax / Rd) * std::exp(-Rmax / Rd);
    for (int i = 0; i < n; ++i) {
        // invert the enclosed fraction by bisection
        const double u = r.next() * cmax;
        double lo = 0.0, hi = Rmax;
        for (int k = 0; k < 40; ++k) {
            const double m = 0.5 * (lo + hi);
            const double f = 1.0 - (1.0 + m / Rd) * std::exp(-m / Rd);
            (f < u ? lo : hi) = m;
        }
        const double R = 0.5 * (lo + hi), ph = 2.0 * M_PI * r.next();
        const double z = h * std::atanh(2.0 * std::fmin(std::fmax(r.next(), 1e-9), 1.0 - 1e-9) - 1.0);
        const double menc = (1.0 - (1.0 + R / Rd) * std::exp(-R / Rd)) / cmax;
        const double vc = std::sqrt(GM * menc / std::fmax(R, 1e-6));
        p[4 * i + 0] = (float)(0.5 + R * std::cos(ph));
        p[4 * i + 1] = (float)(0.5 + R * std::sin(ph));
        p[4 * i + 2] = (float)(0.5 + z);
        v[4 * i + 0] = (float)(-vc * std::sin(ph));
        v[4 * i + 1] = (float)(vc * std::cos(ph));
    }
}

static void cloud(int n, float radius, uint64_t seed, std::vector<float>& p, std::vector<float>& v) {
    p.assign((size_t)n * 4, 0.f);
    v.assign((size_t)n * 4, 0.f);
    Rng r(seed);
    for (int i = 0; i < n; ++i) {
        double x, y, z;
        do {
            x = 2 * r.next() - 1;
            y = 2 * r.next() - 1;
            z = 2 * r.next() - 1;
        } while (x * x + y * y + z * z > 1.0);
        p[4 * i + 0] = (float)(0.5 + radius * x);
        p[4 * i + 1] = (float)(0.5 + radius * y);
        p[4 * i + 2] = (float)(0.5 + radius * z);
    }
}

static SimConfig plain(int n, int G) {
    SimConfig c;
    c.particleCount = n;
    c.gridSize = G;
    c.preset = Preset::Empty;
    c.pressureEnabled = false;
    c.coolingEnabled = false;
    return c;
}

static bool seed_disk(Sim& sim, const SimConfig& c, uint64_t seed) {
    std::vector<float> p, v;
    spiral_disk(c.particleCount, sim.gridSize(), c.gravity, seed, p, v);
    return sim.setParticles(p.data(), v.data(), c.particleCount);
}

// [1] sim_tests.cpp:31-52 -- total grid mass preserved over 50 steps.
static void testMassConservation() {
    std::printf("\n[1] mass conservation\n");
    for (int G : {32, 64, 128}) {
        Sim sim;
        SimConfig cfg = plain(200000, G);
        sim.init(cfg);
        seed_disk(sim, cfg, 1);
        const double m0 = sim.measureTotalGridMass();
        for (int i = 0; i < 50; ++i) sim.step();
        const double m1 = sim.measureTotalGridMass();
        const double rel = std::fabs(m1 - m0) / (m0 > 0 ? m0 : 1.0);
        char buf[160];
        std::snprintf(buf, sizeof(buf), "G=%d m0=%.1f m50=%.1f rel=%.2e", G, m0, m1, rel);
        check(m0 > 1.0 && rel < 1e-3, "deposited mass is conserved", buf);
    }
}

// [2] sim_tests.cpp:56-82 -- stronger gravity clusters more.
static void testGravityResponds() {
    std::printf("\n[2] gravity responds\n");
    auto peakAfter = [](float gravity) {
        Sim sim;
        SimConfig cfg = plain(300000, 128);
        cfg.gravity = gravity;
        sim.init(cfg);
        // the same cold disk (no rotation) in both runs: without gravity it
        // stays put, with gravity it contracts
        std::vector<float> p, v;
        spiral_disk(cfg.particleCount, sim.gridSize(), 2.5f, 2, p, v);
        std::fill(v.begin(), v.end(), 0.f);
        sim.setParticles(p.data(), v.data(), cfg.particleCount);
        for (int i = 0; i < 900; ++i) sim.step();
        return sim.measureMaxDensity();
    };
    const double off = peakAfter(0.0f);
    const double on = peakAfter(2.5f);
    char buf[160];
    std::snprintf(buf, sizeof(buf), "peak(g=0)=%.1f peak(g=2.5)=%.1f ratio=%.2f", off, on, on / (off > 0 ? off : 1));
    check(on > off * 1.5, "gravity on clusters more", buf);
}

// [4] sim_tests.cpp:97-129 -- particle count / grid changes keep the mass.
static void testReconfigure() {
    std::printf("\n[4] reconfigure\n");
    Sim sim;
    SimConfig cfg = plain(100000, 128);
    cfg.gravity = 0.0f;
    sim.init(cfg);
    struct Case {
        int n, g;
    };
    for (Case c : {Case{500000, 64}, Case{100000, 128}, Case{1000000, 128}}) {
        cfg.particleCount = c.n;
        cfg.gridSize = c.g;
        sim.reconfigure(cfg);
        seed_disk(sim, cfg, 3);
        const double mReset = sim.measureTotalGridMass();
        sim.step();
        const double mass = sim.measureTotalGridMass();
        const double rel = std::fabs(mass - c.n) / c.n;
        char buf[200];
        std::snprintf(buf, sizeof(buf), "N=%d G=%d reset=%.0f step=%.0f rel=%.1e", c.n, c.g, mReset, mass, rel);
        check(rel < 1e-3, "mass matches the requested count", buf);
    }
}

// [6] sim_tests.cpp:160-176 -- timings are filled.
static void testTimings() {
    std::printf("\n[6] timings\n");
    Sim sim;
    SimConfig cfg = plain(500000, 128);
    sim.init(cfg);
    seed_disk(sim, cfg, 4);
    for (int i = 0; i < 20; ++i) sim.step();
    const SimTimings t = sim.timings();
    char buf[200];
    std::snprintf(buf, sizeof(buf), "total=%.3f ms dt=%.7f vmax=%.3f", t.totalMs, t.dtUsed, t.maxSpeed);
    check(t.totalMs > 0.0f && t.dtUsed > 0.0f, "step time and dt are filled", buf);
}

// [7] sim_tests.cpp:182-208 -- CFL clamp keeps mass and centroid under strong gravity.
static void testCflClamp() {
    std::printf("\n[7] CFL clamp\n");
    Sim sim;
    SimConfig cfg = plain(300000, 128);
    cfg.gravity = 2.0f;
    sim.init(cfg);
    seed_disk(sim, cfg, 5);
    const double m0 = sim.measureTotalGridMass();
    double cx0, cy0;
    sim.measureCentroid(cx0, cy0);
    for (int i = 0; i < 400; ++i) sim.step();
    const double m1 = sim.measureTotalGridMass();
    double cx1, cy1;
    sim.measureCentroid(cx1, cy1);
    const SimTimings t = sim.timings();
    const double dm = std::fabs(m1 - m0) / (m0 > 0 ? m0 : 1.0);
    const double dc = std::sqrt((cx1 - cx0) * (cx1 - cx0) + (cy1 - cy0) * (cy1 - cy0));
    char buf[220];
    std::snprintf(buf, sizeof(buf), "dm=%.2e dc=%.4f dt=%.2e vmax=%.2f", dm, dc, t.dtUsed, t.maxSpeed);
    check(dm < 1e-2 && dc < 0.05, "mass and centroid held under strong gravity", buf);
}

// [13] sim_tests.cpp:423-458 -- softening acts on the periodic mesh.
static void testPeriodicSoftening() {
    std::printf("\n[13] periodic softening\n");
    auto speedWith = [](float soft) {
        Sim sim;
        SimConfig cfg = plain(200000, 128);
        cfg.boundary = Boundary::Periodic;
        cfg.softeningCells = soft;
        cfg.gravity = 3.0f;
        sim.init(cfg);
        std::vector<float> p, v;
        cloud(200000, 0.18f, 6, p, v);
        sim.setParticles(p.data(), v.data(), 200000);
        for (int i = 0; i < 600; ++i) sim.step();
        return (double)sim.timings().maxSpeed;
    };
    const double sharp = speedWith(0.5f);
    const double blunt = speedWith(6.0f);
    char buf[200];
    std::snprintf(buf, sizeof(buf), "vmax(0.5 cells)=%.6f vmax(6 cells)=%.6f ratio=%.2f", sharp, blunt,
                  blunt > 0 ? sharp / blunt : 0.0);
    check(sharp > blunt * 1.15, "larger softening weakens close forces", buf);
}

// [15] sim_tests.cpp:512-543 -- the same request does not reallocate.
static void testNoReallocOnSameRequest() {
    std::printf("\n[15] no reallocation on the same request\n");
    Sim sim;
    SimConfig cfg = plain(500000, 256);
    sim.init(cfg);
    seed_disk(sim, cfg, 7);
    for (int i = 0; i < 50; ++i) sim.step();
    const double tBefore = sim.simTime();
    for (int i = 0; i < 200; ++i) sim.reconfigure(cfg);
    const double tAfterSame = sim.simTime();
    SimConfig bigger = cfg;
    bigger.particleCount = 700000;
    sim.reconfigure(bigger);
    const double tAfterChange = sim.simTime();
    seed_disk(sim, bigger, 8);
    const int nAfterChange = sim.activeCount();
    char buf[240];
    std::snprintf(buf, sizeof(buf), "t=%.6f -> %.6f (same) -> %.6f (changed), active %d", tBefore, tAfterSame,
                  tAfterChange, nAfterChange);
    check(tBefore > 0.0 && tAfterSame == tBefore && tAfterChange == 0.0 && nAfterChange == 700000,
          "same request kept, changed request reallocated", buf);
}

// State files round-trip (Sim.cu:4985-5120).
static void testStateRoundTrip() {
    std::printf("\n[s] STARDUST state file\n");
    Sim a;
    SimConfig cfg = plain(50000, 64);
    a.init(cfg);
    seed_disk(a, cfg, 9);
    for (int i = 0; i < 5; ++i) a.step();
    const std::string path = "/tmp/stardust_state_test.bin";
    const bool saved = a.saveState(path);
    Sim b;
    b.init(plain(1000, 32));
    const bool loaded = b.loadState(path);
    std::vector<float> pa(4 * 50000), pb(4 * 50000);
    a.getParticles(pa.data(), nullptr, 50000);
    b.getParticles(pb.data(), nullptr, 50000);
    bool same = b.gridSize() == 64 && b.particleCount() == 50000;
    for (size_t i = 0; i < pa.size() && same; ++i) same = pa[i] == pb[i];
    check(saved && loaded && same, "save -> load reproduces the particles", "");
}

// Off-path physics is refused, and refusals are sticky as the reference's g_failed.
static void testRefusal() {
    std::printf("\n[r] off-path features refused\n");
    Sim sim;
    SimConfig cfg = plain(1000, 32);
    cfg.blackHoleEnabled = true;
    sim.init(cfg);
    check(Sim::failed() && !Sim::failMessage().empty(), "black holes refused with a message", Sim::failMessage());
}

int main() {
    std::printf("=== PM gravity core regression checks (C++ Sim over the HIP library) ===\n");
    if (!Sim::deviceAvailable()) {
        std::printf("no HIP device found\n");
        return 2;
    }
    std::printf("GPU: %s\n", Sim::deviceName().c_str());
    testMassConservation();
    testGravityResponds();
    testReconfigure();
    testTimings();
    testCflClamp();
    testPeriodicSoftening();
    testNoReallocOnSameRequest();
    testStateRoundTrip();
    testRefusal();  // last: it leaves the sticky failure set
    std::printf("\n=== result: %d PASS / %d FAIL ===\n", g_pass, g_fail);
    return g_fail == 0 ? 0 : 1;
}
