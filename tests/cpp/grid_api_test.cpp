// C++ boundary test of the HIP backend through the FluidGridHIP adapter (fluidsim_amd/host/).
// Mirrors the contract checks of the reference's tests/FluidGridAPITests.cpp: deterministic exact
// results with and without timings (:83-134), zero-time and zero-iteration steps reported, not thrown
// (:186-229), upload validation, the hydrostatic closed form of tests/FluidGridTests.cpp:140-160,
// checkpoint continuation (tests/FluidWebGPUAPITests.cpp:518-666), the scene frame schedule, and the
// obstacle-mask and multigrid-storage extensions.
// Needs a GPU; run by tests/test_cpp_adapter.py.
#include "fluidsim_amd/host/FluidGridHIP.hpp"

#include <algorithm>
#include <cmath>
#include <iostream>
#include <memory>
#include <stdexcept>

using fluid_hip::FluidGridHIP;
using fluid_hip::FluidMixture;
using fluid_hip::HipContext;

namespace {

void require(bool condition, const char* message)
{
    if (!condition)
        throw std::runtime_error(message);
}

bool sameState(const FluidGridHIP::State& a, const FluidGridHIP::State& b)
{
    return a.liquid_fraction == b.liquid_fraction && a.u == b.u && a.v == b.v && a.pressure == b.pressure &&
           a.liquid_mass == b.liquid_mass && a.gas_mass == b.gas_mass;
}

fluid_hip::SubstanceCatalogue catalogue()
{
    fluid_hip::SubstanceCatalogue c;
    c.add({"water", {0.1f, 0.4f, 0.95f}});
    c.add({"wine", {0.6f, 0.05f, 0.2f}});
    c.add({"air", {0.3f, 0.3f, 0.3f}});
    c.add({"smoke", {0.05f, 0.05f, 0.05f}});
    return c;
}

FluidGridHIP::State configure(FluidGridHIP& grid)
{
    FluidMixture liquid, gas;
    liquid.components = {{0, 100.0}, {1, 300.0}};
    gas.components = {{2, 0.1}, {3, 0.1}};
    grid.configureMaterials(catalogue(), liquid, gas);
    grid.setOrigin({-7.0, 11.0});
    grid.setGravityY(80.0);
    const int width = grid.getWidth(), height = grid.getHeight();
    const double area = grid.getCellSize() * grid.getCellSize();
    FluidGridHIP::State state;
    state.liquid_fraction.resize(width * height);
    state.liquid_mass.assign(2, std::vector<double>(width * height));
    state.gas_mass.assign(2, std::vector<double>(width * height));
    for (int y = 0; y < height; ++y) {
        for (int x = 0; x < width; ++x) {
            const int i = y * width + x;
            const double f = std::clamp(y + 1.0 - (x + height * 0.5), 0.0, 1.0);
            const double red_share = x < width / 4 ? 1.0 : 0.0;
            const double cloud_share = 0.15 + 0.1 * std::sin(0.2 * x);
            state.liquid_fraction[i] = f;
            state.liquid_mass[0][i] = 100.0 * f * area * (1.0 - red_share);
            state.liquid_mass[1][i] = 300.0 * f * area * red_share;
            state.gas_mass[0][i] = 0.1 * (1.0 - f) * area * (1.0 - cloud_share);
            state.gas_mass[1][i] = 0.1 * (1.0 - f) * area * cloud_share;
        }
    }
    grid.upload(state);
    return state;
}

void timedAndUntimedStepsAreIdentical(HipContext& ctx)
{
    FluidGridHIP plain(ctx, 32, 32, 1.0), measured(ctx, 32, 32, 1.0);
    for (auto* g : {&plain, &measured}) {
        g->setTransport(FS_TRANSPORT_SEMI_LAGRANGIAN);
        configure(*g);
    }
    FluidGridHIP::StepTimings timings;
    for (int tick = 0; tick < 8; ++tick) {
        const auto a = plain.step(0.01, 2000, 1e-3);
        const auto b = measured.step(0.01, 2000, 1e-3, &timings);
        require(a.converged && b.converged, "timed comparison did not converge");
        require(a.iterations == b.iterations && a.max_divergence == b.max_divergence, "timings changed the result");
        require(sameState(plain.download(), measured.download()), "timings changed the state");
        require(timings.transport_ms > 0.0 && timings.pressure_ms > 0.0, "missing stage measurements");
    }
    measured.step(0.0, 2000, 1e-3, &timings);
    require(timings.transport_ms == 0.0 && timings.forces_ms == 0.0 && timings.pressure_ms == 0.0,
        "zero-time step retained stale timings");
}

void zeroIterationAndZeroTimeAreReported(HipContext& ctx)
{
    FluidGridHIP grid(ctx, 16, 12, 1.0);
    grid.setTransport(FS_TRANSPORT_NONE);
    configure(grid);
    const auto before = grid.download();
    const auto zero = grid.step(0.0);
    require(!zero.converged && zero.physical_steps == 0, "zero timestep unexpectedly advanced");
    require(sameState(before, grid.download()), "zero timestep changed the state");
    const auto budget = grid.step(0.01, 0, 1e-8);
    require(!budget.converged, "zero-iteration projection unexpectedly converged");
}

void uploadIsValidated(HipContext& ctx)
{
    FluidGridHIP grid(ctx, 8, 6, 1.0);
    bool threw = false;
    try {
        grid.project(0.01);
    } catch (const std::logic_error&) {
        threw = true;
    }
    require(threw, "use before upload was not a logic error");
    auto state = configure(grid);
    state.u.assign(5, 0.0);
    threw = false;
    try {
        grid.upload(state);
    } catch (const std::invalid_argument&) {
        threw = true;
    }
    require(threw, "mismatched upload was not rejected");
    threw = false;
    try {
        grid.setSurfaceTension(-1.0);
    } catch (const std::invalid_argument&) {
        threw = true;
    }
    require(threw, "negative surface tension was not rejected");
    grid.setSurfaceTension(0.07);
    require(grid.getSurfaceTension() == 0.07 && std::isfinite(grid.capillaryStepLimit()),
        "surface tension was not stored");
}

void hydrostaticPressure(HipContext& ctx)
{
    FluidGridHIP grid(ctx, 12, 10, 2.0);
    FluidMixture liquid, gas;
    liquid.components = {{0, 100.0}};
    gas.components = {{1, 0.1}};
    grid.configureMaterials(catalogue(), liquid, gas);
    grid.setGravityY(100.0);
    grid.setTransport(FS_TRANSPORT_NONE);
    FluidGridHIP::State s;
    s.liquid_fraction.assign(120, 1.0);
    s.liquid_mass.assign(1, std::vector<double>(120, 100.0 * 4.0));
    s.gas_mass.assign(1, std::vector<double>(120, 0.0));
    grid.upload(s);
    const auto r = grid.step(1.0 / 120.0, 2000, 1e-8);
    require(r.converged, "hydrostatic projection did not converge");
    const auto out = grid.download();
    for (double v : out.v)
        require(std::abs(v) < 1e-7, "pressure did not cancel gravity");
    const double rise = out.pressure[9 * 12] - out.pressure[0];
    const double expected = 100.0 * 100.0 * 2.0 * 9;
    require(std::abs(rise - expected) < expected * 1e-8, "incorrect hydrostatic pressure gradient");
}

// tests/FluidWebGPUAPITests.cpp:518-666 through the C++ adapter: a restored instance continues
// bit for bit, saving changes nothing, and a corrupt checkpoint is rejected without mutation.
void checkpointContinuation(HipContext& ctx)
{
    FluidGridHIP original(ctx, 33, 25, 2.5), resumed(ctx, 33, 25, 2.5);
    configure(original);
    original.setComponentViscosity(fluid_hip::Phase::Liquid, 1, 40.0);
    for (int tick = 0; tick < 20; ++tick)
        require(original.step(0.004, 2000, 0.01).converged, "checkpoint warm-up failed");
    const auto before = original.download();
    fluid_hip::FluidCheckpoint saved;
    require(original.saveCheckpoint(saved), "saving failed");
    require(sameState(before, original.download()), "saving changed the state");
    require(saved.data.size() == ((3 + 4) * 33 * 25 + 34 * 25 + 33 * 26) * sizeof(double), "payload size");
    require(resumed.loadCheckpoint(saved), "loading failed");
    require(sameState(before, resumed.download()), "loading changed the state");
    require(resumed.getGravityY() == 80.0 && resumed.getOrigin().x == -7.0 &&
                resumed.getLiquidMixture().components.size() == 2 &&
                resumed.getLiquidMixture().components[1].dynamic_viscosity == 40.0,
        "loading lost the configuration");
    for (int tick = 0; tick < 20; ++tick) {
        const auto a = original.step(0.004, 2000, 0.01);
        const auto b = resumed.step(0.004, 2000, 0.01);
        require(a.converged && b.converged && a.iterations == b.iterations &&
                    a.viscosity_iterations == b.viscosity_iterations && a.max_divergence == b.max_divergence,
            "resumed solve differs");
    }
    require(sameState(original.download(), resumed.download()), "resumed trajectory differs");
    fluid_hip::FluidCheckpoint a, b;
    original.saveCheckpoint(a);
    resumed.saveCheckpoint(b);
    require(a.metadata == b.metadata && a.data == b.data, "continued checkpoints differ");
    auto corrupt = a;
    corrupt.metadata.replace(corrupt.metadata.find("\"width\":33"), 10, "\"width\":34");
    bool threw = false;
    try {
        resumed.loadCheckpoint(corrupt);
    } catch (const std::runtime_error&) {
        threw = true;
    }
    fluid_hip::FluidCheckpoint after;
    resumed.saveCheckpoint(after);
    require(threw && after.metadata == b.metadata && after.data == b.data, "corrupt checkpoint was not rejected");
}

// FluidScene.inl:343-393: ceil(frame_time * rate) equal steps, statistics summed over them.
void sceneFrame(HipContext& ctx)
{
    FluidGridHIP framed(ctx, 24, 20, 1.0), stepped(ctx, 24, 20, 1.0);
    configure(framed);
    configure(stepped);
    const auto stats = framed.advanceFrame(0.02, 60, 0.1, true);
    int iterations = 0;
    for (int k = 0; k < 2; ++k) {
        const auto r = stepped.step(0.01, -1, 0.1);
        require(r.converged, "frame reference step failed");
        iterations += r.iterations;
    }
    require(stats.accepted && stats.steps_scheduled == 2 && stats.steps_run == 2 && stats.substeps == 2,
        "frame schedule");
    require(stats.pressure_iterations == iterations && stats.stages.pressure_ms > 0.0, "frame statistics");
    require(sameState(framed.download(), stepped.download()), "frame differs from its explicit steps");
}

// Obstacles and 16-bit multigrid levels through the adapter: solid faces stay closed, every storage mode
// converges, an enclosed pocket is refused, and removing the mask restores the unmasked result.
void obstaclesAndPrecision(HipContext& ctx)
{
    const int W = 160, H = 120;
    std::vector<unsigned char> mask((size_t)W * H, 0);
    for (int y = 40; y < 80; ++y)
        for (int x = 60; x < 70; ++x)
            mask[(size_t)y * W + x] = 1;
    std::vector<double> u_after;
    for (fs_mg_precision mode : {FS_MG_FP32, FS_MG_FP16, FS_MG_BF16, FS_MG_BF16_VECTORS}) {
        FluidGridHIP grid(ctx, W, H, 1.0 / W);
        configure(grid);
        grid.setSolidMask(mask);
        grid.setMultigridPrecision(mode);
        require(grid.solidCells() == 400 && grid.getMultigridPrecision() == mode, "configuration");
        grid.addVortex({-6.5, 11.35}, 0.1, 1.0);
        const auto r = grid.project(0.01, 500, 1e-6);
        require(r.converged, "masked projection did not converge");
        const auto s = grid.download();
        for (int y = 40; y < 80; ++y)
            for (int x = 60; x <= 70; ++x)
                require(s.u[(size_t)y * (W + 1) + x] == 0.0, "a solid face carries velocity");
    }
    FluidGridHIP plain(ctx, W, H, 1.0 / W), toggled(ctx, W, H, 1.0 / W);
    configure(plain);
    configure(toggled);
    toggled.setSolidMask(mask);
    toggled.setSolidMask({});
    plain.addVortex({-6.5, 11.35}, 0.1, 1.0);
    toggled.addVortex({-6.5, 11.35}, 0.1, 1.0);
    require(plain.project(0.01, 500, 1e-6).iterations == toggled.project(0.01, 500, 1e-6).iterations,
        "removing the mask changed the solve");
    require(sameState(plain.download(), toggled.download()), "removing the mask changed the state");
    std::vector<unsigned char> pocket((size_t)W * H, 0);
    for (int x = 0; x < W; ++x)
        pocket[(size_t)60 * W + x] = 1;
    bool threw = false;
    try {
        plain.setSolidMask(pocket);
    } catch (const std::invalid_argument&) {
        threw = true;
    }
    require(threw && plain.solidCells() == 0, "a mask splitting the fluid was accepted");
}

void batchedSolves(HipContext& ctx)
{
    // grids of 80 x 60 (4800 cells) and 160 x 120 (19200): the split sends the larger through its own
    // solve; every result matches the grid's single projection to the solve tolerance
    const int sizes[2][2] = {{80, 60}, {160, 120}};
    std::vector<std::unique_ptr<FluidGridHIP>> grids, singles;
    std::vector<FluidGridHIP*> batch;
    for (const auto& wh : sizes) {
        for (auto* list : {&grids, &singles}) {
            list->push_back(std::make_unique<FluidGridHIP>(ctx, wh[0], wh[1], 1.0 / wh[0]));
            configure(*list->back());
            list->back()->addVortex({-6.5, 11.35}, 0.1, 1.0);
        }
        batch.push_back(grids.back().get());
    }
    for (long long split : {0LL, 10000LL}) {
        ctx.setBatchSplit(split);
        const auto results = batchProject(batch, 0.01, 500, {1e-6, 1e-6});
        for (size_t k = 0; k < batch.size(); ++k)
            require(results[k].converged && results[k].max_divergence <= 1e-6, "batched projection failed");
    }
    ctx.setBatchSplit(0);
    for (size_t k = 0; k < batch.size(); ++k) {
        const auto r = singles[k]->project(0.01, 500, 1e-6);
        require(r.converged, "single projection failed");
    }
    bool threw = false;
    try {
        batchProject(batch, 0.01, 500, {1e-6});
    } catch (const std::invalid_argument&) {
        threw = true;
    }
    require(threw, "a tolerance list of the wrong length was accepted");
    const auto its = batchApplyViscosity(batch, 0.01);
    require(its.size() == batch.size(), "batched viscosity result count");
}

// writeColorPixels and commitEdits through the adapter (FluidGridBase.h:139-142): caller storage is
// resized and filled, pure cells show their substance colour, edits replace material only, the last edit
// of a cell wins, a bad batch changes nothing, and velocity edits are refused as on UniformCPU.
void colorsAndEdits(HipContext& ctx)
{
    FluidGridHIP grid(ctx, 24, 16, 0.5);
    const FluidGridHIP::State before = configure(grid);
    std::vector<float> pixels(3, -1.0f);
    grid.writeColorPixels(pixels);
    require(pixels.size() == 24 * 16 * 4, "writeColorPixels resizes the caller's storage");
    auto edit = grid.edit();
    edit.setCell(3, 2, fluid_hip::Phase::Liquid, 1);   // wine
    edit.setCell(5, 9, fluid_hip::Phase::Gas, 3);      // smoke
    edit.setCell(3, 2, fluid_hip::Phase::Liquid, 0);   // water: the last edit of (3, 2) wins
    edit.commit();
    require(edit.empty(), "a successful commit clears the session");
    const FluidGridHIP::State after = grid.download();
    const int a = 2 * 24 + 3, b = 9 * 24 + 5;
    require(after.liquid_fraction[a] == 1.0 && after.liquid_fraction[b] == 0.0, "edited fractions");
    require(after.liquid_mass[0][a] == 100.0 * 0.25 && after.liquid_mass[1][a] == 0.0, "edited liquid masses");
    require(after.gas_mass[1][b] == 0.1 * 0.25 && after.gas_mass[0][b] == 0.0 && after.liquid_mass[0][b] == 0.0,
        "edited gas masses");
    // the uploaded state left u and v empty, which means zero faces
    require(before.u.empty() && before.v.empty(), "the configured state has no velocities");
    require(after.u == std::vector<double>(25 * 16, 0.0) && after.v == std::vector<double>(24 * 17, 0.0),
        "edits keep the face velocities");
    grid.writeColorPixels(pixels, true);
    require(pixels[4 * a] == 0.1f && pixels[4 * a + 1] == 0.4f && pixels[4 * a + 2] == 0.95f && pixels[4 * a + 3] == 1.0f,
        "a pure water cell shows the water colour");
    require(pixels[4 * b] == 0.05f && pixels[4 * b + 3] == 0.0f, "a pure smoke cell shows the smoke colour");
    bool threw = false;
    try {
        edit.setCell(0, 0, fluid_hip::Phase::Liquid, 0);
        edit.setCell(24, 0, fluid_hip::Phase::Liquid, 0);
        edit.commit();
    } catch (const std::out_of_range&) {
        threw = true;
    }
    require(threw && grid.download().liquid_mass == after.liquid_mass, "an edit outside the grid rejects the batch");
    edit.clear();
    threw = false;
    try {
        edit.setCell(1, 1, fluid_hip::Phase::Gas, 0);   // water is not a gas component
        edit.commit();
    } catch (const std::invalid_argument&) {
        threw = true;
    }
    require(threw, "an unconfigured substance is rejected");
    edit.clear();
    threw = false;
    try {
        edit.setCell(1, 1, fluid_hip::Phase::Liquid, 0);
        edit.commit({1.0, 0.0});
    } catch (const std::invalid_argument&) {
        threw = true;
    }
    require(threw, "velocity edits require uniform WebGPU");
}

}  // namespace

int main()
{
    HipContext ctx(0);
    const struct {
        const char* name;
        void (*run)(HipContext&);
    } tests[] = {{"timed and untimed steps are identical", timedAndUntimedStepsAreIdentical},
        {"zero time and zero iterations are reported", zeroIterationAndZeroTimeAreReported},
        {"upload validation", uploadIsValidated}, {"hydrostatic pressure", hydrostaticPressure},
        {"checkpoint continuation", checkpointContinuation}, {"scene frame", sceneFrame},
        {"obstacles and multigrid precision", obstaclesAndPrecision},
        {"batched solves and the batch split", batchedSolves}, {"colours and edits", colorsAndEdits}};
    int failures = 0;
    for (const auto& t : tests) {
        try {
            t.run(ctx);
            std::cout << "PASS " << t.name << '\n';
        } catch (const std::exception& e) {
            ++failures;
            std::cerr << "FAIL " << t.name << ": " << e.what() << '\n';
        }
    }
    return failures == 0 ? 0 : 1;
}
