// host_asan_check.cpp -- the host code paths of libphc.so and of the drop-in
// headers that need no device, run under AddressSanitizer (host side only;
// recipe: proxy-geos-hc_amd/build.py sanitize_host).  Setup routines at every
// half width, order and edge size, the validation paths of the create calls
// (which fail before any device work), the per-XCD schedule tables, the
// colour lists, the drop-in argv parser and mesh helpers, and the Array2D
// binding rules against a host-only column store.  Prints "HOST ASAN OK".
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "FDTD.hpp"
#include "SEM.hpp"

static int failures = 0;
#define CHECK(c)                                                                                   \
    do {                                                                                           \
        if (!(c)) {                                                                                \
            std::printf("check failed: %s (line %d)\n", #c, __LINE__);                            \
            ++failures;                                                                            \
        }                                                                                          \
    } while (0)

// A column store that keeps the "device" columns in host vectors.
struct HostStore : ColumnStore<float> {
    std::vector<float> col[2];
    int pulls = 0, pushes = 0;
    explicit HostStore(size_t n) { col[0].assign(n, 1.f); col[1].assign(n, 2.f); }
    void pull(int c, float *h) override { ++pulls; std::memcpy(h, col[c].data(), col[c].size() * 4); }
    void push(int c, const float *h) override { ++pushes; std::memcpy(col[c].data(), h, col[c].size() * 4); }
};

static void setup_paths()
{
    for (int L = 1; L <= 6; ++L) {
        std::vector<float> c(L + 1);
        CHECK(phc_fd_init_coef(L, 10.f, c.data()) == PHC_OK);
        const double c0 = phc_fd_coef0(c.data(), L, c.data(), L, c.data(), L);
        CHECK(c0 < 0);
        CHECK(phc_fd_compute_dt(4500.f, c.data(), L, c.data(), L, c.data(), L, 0.8f) > 0.f);
    }
    CHECK(phc_fd_init_coef(7, 10.f, nullptr) != PHC_OK);
    std::vector<float> st(777);
    for (int order = 0; order <= 2; ++order)
        for (int ev : {PHC_EXP_DOUBLE, PHC_EXP_FLOAT})
            CHECK(phc_source_term((int)st.size(), 0.001f, 10.f, order, ev, st.data()) == PHC_OK);
    for (int n : {1, 2, 7, 33}) {
        std::vector<float> vp((size_t)n * (n + 1) * (n + 2));
        CHECK(phc_fd_model_two_layer(n, n + 1, n + 2, 1500.f, 0.001f, vp.data()) == PHC_OK);
        std::vector<int> inter(5);
        const int nl = n > 6 ? 6 : 1;
        CHECK(phc_fd_model_layered(n, n + 1, n + 2, nl, 1500.f, 4500.f, 7, 0.001f, inter.data(), vp.data()) ==
              PHC_OK);
        std::vector<float> prof(n + 2);
        const int nd = n / 3;
        if (nd > 0) CHECK(phc_fd_pml_profile(n, nd, phc_fd_pml_scale(0.001f, 4500.f, nd, 10.f), prof.data()) == PHC_OK);
    }
    std::vector<float> g((size_t)7 * 5 * 6);
    CHECK(phc_sem_model_gaussian(7, 5, 6, 2.0, 2000.f, 4000.f, 3, g.data()) == PHC_OK);
    for (int ex : {1, 2, 5, 9})
        for (int ey : {0, 1, 4}) {
            int maxc = 0;
            const int nc = phc_sem_color_count(ex, ey, 3, &maxc);
            std::vector<int> counts(nc), lists((size_t)nc * maxc);
            CHECK(phc_sem_color_elements(ex, ey, 3, counts.data(), lists.data()) == PHC_OK);
            CHECK(phc_sem_element_from_point(ex, ey, 3, 100.f, 100.f, 100.f, 50.f, 50.f, 50.f) >= 0);
        }
    for (int planes : {1, 5, 49, 399})
        for (int cols : {1, 7, 175, 300})
            for (int split = 1; split <= 4; ++split) {
                int n = 0;
                CHECK(phc_sem_schedule_table(planes, cols, 256, split, nullptr, &n) == PHC_OK);
                std::vector<int> tab((size_t)3 * n);
                CHECK(phc_sem_schedule_table(planes, cols, 256, split, tab.data(), &n) == PHC_OK);
            }
}

static void validation_paths()
{
    phc_fd_params p;
    std::memset(&p, 0, sizeof(p));
    phc_fd *h = nullptr;
    CHECK(phc_fd_create(&p, &h) != PHC_OK);  // empty grid
    std::vector<float> c(5), vp(8 * 8 * 8, 0.1f);
    phc_fd_init_coef(4, 10.f, c.data());
    p.nx = p.ny = p.nz = 8; p.lx = 4; p.ly = 9; p.lz = 4;
    p.coefx = p.coefy = p.coefz = c.data(); p.vp = vp.data();
    CHECK(phc_fd_create(&p, &h) != PHC_OK);  // half width 9
    p.ly = 4; p.xs = 8;
    CHECK(phc_fd_create(&p, &h) != PHC_OK);  // source outside
    p.xs = 4; vp[17] = -1.f;
    CHECK(phc_fd_create(&p, &h) != PHC_OK);  // negative vp
    vp[17] = 0.1f; p.k_remainder = 3;
    CHECK(phc_fd_create(&p, &h) != PHC_OK);  // bad mode
    phc_sem_params q;
    std::memset(&q, 0, sizeof(q));
    phc_sem *s = nullptr;
    CHECK(phc_sem_create(&q, &s) != PHC_OK);
    phc_sem2d *s2 = nullptr;
    CHECK(phc_sem2d_create(&q, &s2) != PHC_OK);
    phc_batch_params b;
    std::memset(&b, 0, sizeof(b));
    phc_batch *bh = nullptr;
    CHECK(phc_batch_create(&b, &bh) != PHC_OK);
    CHECK(std::strlen(phc_last_error()) > 0);
}

static void dropin_paths()
{
    const char *argv[] = {"fd", "-nx", "20", "-ny", "18", "-lx", "4", "-ly", "3", "-lz", "2", "-xs", "10",
                          "-usePML", "-fp32", "-nsteps", "10", "-f0", "40"};
    FDTDGRIDS g;
    FDTDInit init;
    init.init_geometry((int)(sizeof(argv) / sizeof(argv[0])), const_cast<char **>(argv), g);
    CHECK(g.nx == 20 && g.ny == 18 && g.nz == 20 && g.lx == 4 && g.ly == 3 && g.lz == 2 && g.xs == 10);
    FDTDMODELS m;
    init.init_coefficients(g, m);
    init.init_source(m);
    init.init_models(g, m);
    init.defineSpongeBoundary(g, m.spongeArray);
    CHECK(m.coefy.size() == 4 && m.RHSTerm.size() == init.nSamples && m.usePML);
    CHECK(m.spongeArray(0) > 0.f);
    const char *bad[] = {"fd", "-lx", "7"};
    bool threw = false;
    try {
        FDTDInit i2;
        FDTDGRIDS g2;
        i2.init_geometry(3, const_cast<char **>(bad), g2);
    } catch (const std::exception &) {
        threw = true;
    }
    CHECK(threw);
    SEMmesh mesh(5, 4, 3, 100.f, 80.f, 60.f, 3);
    arrayInt nodes;
    mesh.globalNodesList(mesh.getNumberOfElements(), nodes);
    vectorInt inner;
    CHECK(mesh.getListOfInteriorNodes(mesh.getNumberOfInteriorNodes(), inner) == mesh.getNumberOfInteriorNodes());
    vectorReal model;
    mesh.getModel(mesh.getNumberOfElements(), model);
    SolverUtils u;
    CHECK(u.computeSourceTerm(50, 0.001f, 10.f, 1).size() == 50);

    // Array2D binding rules
    const int rows = 37;
    arrayReal a = allocateArray2D<arrayReal>(rows, 2, "pn");
    a.fill(0.5f);
    {
        HostStore hs(rows);
        a.bind(&hs);
        CHECK(hs.pushes == 0);          // untouched host copy: nothing sent
        a.device_changed(0);
        const arrayReal &ca = a;
        CHECK(ca(3, 0) == 1.f && hs.pulls == 1);  // read after a device step refreshes
        a(4, 1) = 9.f;                  // host write, sent before the next step
        a.push_host_changes();
        CHECK(hs.pushes == 1 && hs.col[1][4] == 9.f && hs.col[1][5] == 0.5f);
        a.device_changed(0);
        a.device_changed(1);
    }
    // the store is gone with columns still newer there: reading them throws
    threw = false;
    try {
        const arrayReal &ca = a;
        (void)ca(0, 0);
    } catch (const std::exception &) {
        threw = true;
    }
    CHECK(threw);
    arrayReal copy = allocateArray2D<arrayReal>(3, 2, "c");
    copy(1, 1) = 2.f;
    arrayReal copy2 = copy;
    CHECK(copy2(1, 1) == 2.f);
}

int main()
{
    setup_paths();
    validation_paths();
    dropin_paths();
    std::printf(failures ? "HOST ASAN FAILED (%d)\n" : "HOST ASAN OK\n", failures);
    return failures ? 1 : 0;
}
