// FluidGridHIP.hpp -- C++ host adapter: the reference backend interface over the HIP C ABI.
//
// Implements the method set of fluid::FluidGridBase (FluidSim/FluidGridBase.h:20-147) with the same
// names, signatures and error behaviour, forwarding every call to libfluidsim_hip.so
// (include/fluidsim_hip.h).  The value types mirror FluidSim/FluidGridTypes.h:10-79 and
// FluidSim/FluidMixture.h:14-39 field for field.  Inside the reference tree the same class derives
// from fluid::FluidGridBase and is selected in FluidGrid::selectBackend (FluidSim/FluidGrid.cpp:37-90).
//
// Error mapping (FluidGridBase.h:48-116): FS_E_INVALID_ARGUMENT -> std::invalid_argument,
// FS_E_LOGIC -> std::logic_error, FS_E_UNSUPPORTED -> std::invalid_argument (the base class's
// "requires uniform WebGPU" contract), device/runtime failures -> std::runtime_error.
// Non-convergence is returned in ProjectionResult, never thrown.
#pragma once

#include "../../include/fluidsim_hip.h"

#include <algorithm>
#include <array>
#include <cstddef>
#include <cstdint>
#include <limits>
#include <stdexcept>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

namespace fluid_hip {

template <typename Scalar> struct FluidGridState {
    std::vector<Scalar> liquid_fraction;
    std::vector<Scalar> u;
    std::vector<Scalar> v;
    std::vector<Scalar> pressure;
    std::vector<std::vector<Scalar>> liquid_mass;
    std::vector<std::vector<Scalar>> gas_mass;
};

struct FluidStepTimings {
    double transport_ms = 0;
    double viscosity_ms = 0;
    double forces_ms = 0;
    double pressure_ms = 0;
    double selection_ms = 0;
    double regrid_ms = 0;
};

struct FluidProjectionResult {
    int iterations = 0;
    double max_divergence = std::numeric_limits<double>::infinity();
    bool converged = false;
    int physical_steps = 1;
    int transport_steps = 0;
    bool relaxed = false;
    double max_volume_error = 0.0;
    double rms_volume_error = 0.0;
    int viscosity_iterations = 0;
    bool precision_limited = false;

    bool accepted() const { return converged || relaxed || precision_limited; }
};

enum class Phase { Liquid, Gas };

using SubstanceId = std::uint32_t;

// Substance and SubstanceCatalogue (FluidSim/SubstanceCatalogue.h:14-59): the colours writeColorPixels
// reads; SubstanceId is a substance's index in the catalogue.  Chemistry (elements, composition) is not
// used on this path and not mirrored.
struct Substance {
    std::string name;
    std::array<float, 3> color{1.0f, 1.0f, 1.0f};
};

class SubstanceCatalogue {
public:
    SubstanceId add(Substance substance)
    {
        definitions_.push_back(std::move(substance));
        return static_cast<SubstanceId>(definitions_.size() - 1);
    }
    const Substance& operator[](SubstanceId id) const { return definitions_.at(id); }
    std::size_t size() const { return definitions_.size(); }

private:
    std::vector<Substance> definitions_;
};

struct MixtureComponent {
    SubstanceId substance = 0;
    double density = 0.0;
    double dynamic_viscosity = 0.0;
};

struct FluidMixture {
    std::vector<MixtureComponent> components;
    bool hasViscosity() const
    {
        for (const auto& c : components)
            if (c.dynamic_viscosity != 0.0)
                return true;
        return false;
    }
};

// FluidCheckpoint (FluidSim/FluidCheckpoint.h:9-15) with the metadata kept as JSON text.
struct FluidCheckpoint {
    std::string metadata;
    std::vector<std::byte> data;
};

// FluidScene::FrameStatistics (FluidSim/FluidScene.h:162-178) plus the frame's schedule and outcome.
struct FrameStatistics {
    double fluid_ms = 0;
    FluidStepTimings stages;
    int substeps = 0;
    int transport_steps = 0;
    int pressure_iterations = 0;
    int viscosity_iterations = 0;
    double projection_tolerance = 0;
    double max_volume_error = 0;
    double rms_volume_error = 0;
    int steps_scheduled = 0;
    int steps_run = 0;
    bool accepted = false;
};

struct DVec2 {
    double x = 0.0, y = 0.0;
};

// Staged material edits: detail::CellEdit / ResolvedCellEdit / CellEditBatch (FluidSim/FluidGridEdit.h:14-95).
struct CellEdit {
    int x;
    int y;
    Phase phase;
    SubstanceId substance;
};

struct CellEditBatch {
    std::vector<CellEdit> changes;
    std::vector<fs_cell_edit> cells;   // resolved: component (liquids, then gases) and mass
    bool has_velocity = false;
    DVec2 velocity;

    // Resolves every change to a mixture component and its full-cell mass (density * cell volume), checking
    // the whole batch before anything is written: std::out_of_range for a cell outside the grid,
    // std::invalid_argument for a substance the phase does not hold.  Cells end in raster order with the
    // last change of a cell kept.
    void prepare(int width, int height, double cell_volume, const FluidMixture& liquid, const FluidMixture& gas)
    {
        cells.clear();
        cells.reserve(changes.size());
        int component = 0;
        double mass = 0.0;
        for (std::size_t i = 0; i < changes.size(); ++i) {
            const CellEdit& c = changes[i];
            if (c.x < 0 || c.x >= width || c.y < 0 || c.y >= height)
                throw std::out_of_range("Fluid grid edit cell is outside the grid");
            if (i == 0 || c.phase != changes[i - 1].phase || c.substance != changes[i - 1].substance) {
                const FluidMixture& mixture = c.phase == Phase::Liquid ? liquid : gas;
                std::size_t k = 0;
                while (k < mixture.components.size() && mixture.components[k].substance != c.substance)
                    ++k;
                if (k == mixture.components.size())
                    throw std::invalid_argument("Fluid grid edit substance is not configured in the selected phase");
                component = (int)(k + (c.phase == Phase::Liquid ? 0 : liquid.components.size()));
                mass = mixture.components[k].density * cell_volume;
            }
            cells.push_back(fs_cell_edit{c.x, c.y, component, mass});
        }
        auto raster = [](const fs_cell_edit& a, const fs_cell_edit& b) {
            return a.y < b.y || (a.y == b.y && a.x < b.x);
        };
        std::stable_sort(cells.begin(), cells.end(), raster);
        std::size_t kept = 0;
        for (std::size_t i = 0; i < cells.size(); ++i) {
            if (kept > 0 && cells[kept - 1].x == cells[i].x && cells[kept - 1].y == cells[i].y)
                cells[kept - 1] = cells[i];
            else
                cells[kept++] = cells[i];
        }
        cells.resize(kept);
    }
};

template <class Grid> class FluidGridEdit;

inline void check(fs_status status)
{
    if (status == FS_OK)
        return;
    const std::string message = fs_last_error();
    switch (status) {
    case FS_E_INVALID_ARGUMENT:
    case FS_E_UNSUPPORTED: throw std::invalid_argument(message);
    case FS_E_LOGIC: throw std::logic_error(message);
    default: throw std::runtime_error(message);
    }
}

// One HIP device and one stream shared by the grids created on it.
class HipContext {
public:
    explicit HipContext(int device = 0) { check(fs_context_create(device, &ctx_)); }
    ~HipContext() { fs_context_destroy(ctx_); }
    HipContext(const HipContext&) = delete;
    HipContext& operator=(const HipContext&) = delete;
    fs_context* get() const { return ctx_; }
    // fs_context_set_batch_split: batched solves send grids above `solo_cells` cells through their
    // single-grid solve beside the persistent launch (0: batch every grid).
    void setBatchSplit(long long solo_cells) { check(fs_context_set_batch_split(ctx_, solo_cells)); }

private:
    fs_context* ctx_ = nullptr;
};

class FluidGridHIP {
public:
    using Scalar = double;
    using State = FluidGridState<Scalar>;
    using ProjectionResult = FluidProjectionResult;
    using StepTimings = FluidStepTimings;

    enum class PressurePreconditioner { Diagonal, Multigrid };

    FluidGridHIP(HipContext& context, int width, int height, double cell_size)
        : width_(width), height_(height), cell_size_(cell_size)
    {
        check(fs_grid_create(context.get(), width, height, cell_size, &grid_));
    }
    ~FluidGridHIP() { destroy(); }
    FluidGridHIP(const FluidGridHIP&) = delete;
    FluidGridHIP& operator=(const FluidGridHIP&) = delete;

    int getWidth() const { return width_; }
    int getHeight() const { return height_; }
    double getCellSize() const { return cell_size_; }
    DVec2 getOrigin() const
    {
        DVec2 o;
        check(fs_get_origin(grid_, &o.x, &o.y));
        return o;
    }
    void setOrigin(DVec2 origin) { check(fs_set_origin(grid_, origin.x, origin.y)); }
    bool requiresRenderThread() const { return false; }

    // configureMaterials(SubstanceCatalogue, FluidMixture, FluidMixture) (FluidGridBase.h:42): the
    // catalogue supplies the component colours of writeColorPixels.
    void configureMaterials(SubstanceCatalogue substances, FluidMixture liquid, FluidMixture gas)
    {
        std::vector<double> ld, lv, gd, gv;
        for (const auto& c : liquid.components) {
            ld.push_back(c.density);
            lv.push_back(c.dynamic_viscosity);
        }
        for (const auto& c : gas.components) {
            gd.push_back(c.density);
            gv.push_back(c.dynamic_viscosity);
        }
        check(fs_configure_materials(grid_, (int)ld.size(), ld.data(), lv.data(), (int)gd.size(), gd.data(),
            gv.data()));
        substances_ = std::move(substances);
        liquid_ = std::move(liquid);
        gas_ = std::move(gas);
    }
    const SubstanceCatalogue& getSubstances() const { return substances_; }
    const FluidMixture& getLiquidMixture() const { return liquid_; }
    const FluidMixture& getGasMixture() const { return gas_; }
    bool hasViscosity() const { return liquid_.hasViscosity() || gas_.hasViscosity(); }
    // FluidGridBase.h:48-52; out-of-range components throw std::invalid_argument.
    void setComponentViscosity(Phase phase, size_t component, double coefficient)
    {
        check(fs_set_component_viscosity(grid_, phase == Phase::Liquid ? 0 : 1, (int)component, coefficient));
        auto& mixture = phase == Phase::Liquid ? liquid_ : gas_;
        mixture.components.at(component).dynamic_viscosity = coefficient;
    }
    // FluidGridGPU::applyViscosity (FluidGridGPU.cpp:1039-1046): CG iterations dispatched.
    int applyViscosity(double dt)
    {
        int iterations = 0;
        check(fs_apply_viscosity(grid_, dt, &iterations));
        return iterations;
    }

    double getGravityY() const
    {
        double g = 0.0;
        check(fs_get_gravity(grid_, &g));
        return g;
    }
    void setGravityY(double acceleration) { check(fs_set_gravity(grid_, acceleration)); }
    double getSurfaceTension() const { return surface_tension_; }
    void setSurfaceTension(double coefficient)
    {
        check(fs_set_surface_tension(grid_, coefficient));
        surface_tension_ = coefficient;
    }
    double capillaryStepLimit() const
    {
        double v = 0.0;
        check(fs_capillary_step_limit(grid_, &v));
        return v;
    }

    void setPressurePreconditioner(PressurePreconditioner kind)
    {
        check(fs_set_preconditioner(grid_, kind == PressurePreconditioner::Diagonal ? FS_PRECOND_DIAGONAL
                                                                                   : FS_PRECOND_MULTIGRID));
    }
    void setTransport(fs_transport mode) { check(fs_set_transport(grid_, mode)); }

    // Solid obstacles (no reference counterpart): W*H cells, nonzero = solid; an empty vector removes them.
    void setSolidMask(const std::vector<unsigned char>& mask)
    {
        if (!mask.empty() && mask.size() != static_cast<size_t>(width_) * height_)
            throw std::invalid_argument("solid mask dimensions do not match the grid");
        check(fs_set_solid_mask(grid_, mask.empty() ? nullptr : mask.data()));
    }
    // Multigrid level storage (no reference counterpart): FS_MG_FP32, FS_MG_FP16 or FS_MG_BF16.
    void setMultigridPrecision(fs_mg_precision mode) { check(fs_set_multigrid_precision(grid_, mode)); }
    // step/project tolerances as relative residuals (fs_set_relative_tolerance)
    void setRelativeTolerance(bool enabled) { check(fs_set_relative_tolerance(grid_, enabled ? 1 : 0)); }
    // Smoother of the multigrid levels above 4096 cells and sweeps of the smaller ones
    // (fs_set_pressure_smoother; build's own setting).
    void setPressureSmoother(fs_pressure_smoother kind, int sweeps, int coarse_sweeps = 4)
    {
        check(fs_set_pressure_smoother(grid_, kind, sweeps, coarse_sweeps));
    }
    // Viscosity smoother (fs_set_viscosity_smoother): 0 the reference's V(2,2) damped Jacobi, 2/4/6 Chebyshev sweeps.
    void setViscositySmoother(int chebyshev_sweeps) { check(fs_set_viscosity_smoother(grid_, chebyshev_sweeps)); }
    // Keeps the fastest converging viscosity smoother for a solve of the current state; returns its sweeps.
    int tuneViscositySmoother(double dt)
    {
        int n = 0;
        check(fs_tune_viscosity_smoother(grid_, dt, &n, nullptr));
        return n;
    }
    // Keeps the fastest converging smoother for a projection of the current state (fs_tune_pressure_smoother);
    // returns (kind, sweeps, coarse sweeps).
    std::tuple<fs_pressure_smoother, int, int> tunePressureSmoother(double dt, int max_iterations,
        double divergence_tolerance)
    {
        fs_pressure_smoother kind = FS_SMOOTHER_RED_BLACK;
        int sweeps = 4, coarse = 4;
        check(fs_tune_pressure_smoother(grid_, dt, max_iterations, divergence_tolerance, &kind, &sweeps, &coarse,
            nullptr));
        return {kind, sweeps, coarse};
    }
    fs_mg_precision getMultigridPrecision() const
    {
        fs_mg_precision m = FS_MG_FP32;
        check(fs_get_multigrid_precision(grid_, &m));
        return m;
    }
    long long solidCells() const
    {
        long long n = 0;
        check(fs_solid_cells(grid_, &n));
        return n;
    }

    // FluidGridCPU::upload (FluidGridCPU.cpp:198-251): empty vectors mean zero; sizes must match.
    void upload(const State& state)
    {
        const size_t n = (size_t)width_ * height_;
        const size_t nu = (size_t)(width_ + 1) * height_, nv = (size_t)width_ * (height_ + 1);
        auto field = [&](const std::vector<double>& v, size_t count) -> const double* {
            if (v.empty())
                return nullptr;
            if (v.size() != count)
                throw std::invalid_argument("FluidGridHIP upload dimensions do not match the grid");
            return v.data();
        };
        auto components = [&](const std::vector<std::vector<double>>& src, size_t k, std::vector<double>& flat)
            -> const double* {
            if (src.empty())
                return nullptr;
            if (src.size() != k)
                throw std::invalid_argument("FluidGridHIP upload component count does not match the mixture");
            flat.assign(k * n, 0.0);
            for (size_t c = 0; c < k; ++c) {
                const double* p = field(src[c], n);
                if (p)
                    std::copy(p, p + n, flat.begin() + c * n);
            }
            return flat.data();
        };
        std::vector<double> lm, gm;
        check(fs_upload(grid_, field(state.liquid_fraction, n), field(state.u, nu), field(state.v, nv),
            field(state.pressure, n), components(state.liquid_mass, liquid_.components.size(), lm),
            components(state.gas_mass, gas_.components.size(), gm)));
    }

    State download() const
    {
        const size_t n = (size_t)width_ * height_;
        const size_t kl = liquid_.components.size(), kg = gas_.components.size();
        State s;
        s.liquid_fraction.resize(n);
        s.u.resize((size_t)(width_ + 1) * height_);
        s.v.resize((size_t)width_ * (height_ + 1));
        s.pressure.resize(n);
        std::vector<double> lm(kl * n), gm(kg * n);
        check(fs_download(grid_, s.liquid_fraction.data(), s.u.data(), s.v.data(), s.pressure.data(),
            kl ? lm.data() : nullptr, kg ? gm.data() : nullptr));
        for (size_t c = 0; c < kl; ++c)
            s.liquid_mass.emplace_back(lm.begin() + c * n, lm.begin() + (c + 1) * n);
        for (size_t c = 0; c < kg; ++c)
            s.gas_mass.emplace_back(gm.begin() + c * n, gm.begin() + (c + 1) * n);
        return s;
    }

    void invalidateCaches() { check(fs_invalidate_caches(grid_)); }

    // writeColorPixels (FluidGridBase.h:139, FluidGridCPU.cpp:1826-1860): RGBA per cell into caller storage.
    void writeColorPixels(std::vector<float>& pixels, bool liquid_fraction_alpha = false) const
    {
        std::vector<float> rgb;
        for (const FluidMixture* m : {&liquid_, &gas_})
            for (const auto& c : m->components) {
                const Substance& sub = substances_[c.substance];
                rgb.insert(rgb.end(), sub.color.begin(), sub.color.end());
            }
        pixels.resize((size_t)width_ * height_ * 4);
        check(fs_color(grid_, rgb.empty() ? nullptr : rgb.data(), liquid_fraction_alpha ? 1 : 0, pixels.data()));
    }

    // commitEdits (FluidGridBase.h:142, FluidGridCPU.cpp:253-284).  Velocity edits are a uniform-WebGPU
    // feature: std::invalid_argument, as UniformCPU.
    void commitEdits(CellEditBatch& batch)
    {
        if (batch.has_velocity)
            throw std::invalid_argument("Material edit velocity requires uniform WebGPU");
        batch.prepare(width_, height_, cell_size_ * cell_size_, liquid_, gas_);
        check(fs_commit_edits(grid_, (int)batch.cells.size(), batch.cells.empty() ? nullptr : batch.cells.data(),
            nullptr));
    }

    FluidGridEdit<FluidGridHIP> edit();

    void destroy()
    {
        if (grid_)
            fs_grid_destroy(grid_);
        grid_ = nullptr;
    }

    ProjectionResult step(double dt, int max_iterations = 200, double divergence_tolerance = 1e-6,
        StepTimings* timings = nullptr)
    {
        fs_projection_result r{};
        fs_step_timings t{};
        check(fs_step(grid_, dt, max_iterations, divergence_tolerance, timings ? &t : nullptr, &r));
        if (timings)
            *timings = {t.transport_ms, t.viscosity_ms, t.forces_ms, t.pressure_ms, t.selection_ms, t.regrid_ms};
        return convert(r);
    }

    ProjectionResult project(double dt, int max_iterations = 200, double divergence_tolerance = 1e-6)
    {
        fs_projection_result r{};
        check(fs_project(grid_, dt, max_iterations, divergence_tolerance, &r));
        return convert(r);
    }

    void applyGravity(double dt) { check(fs_apply_gravity(grid_, dt)); }
    void addVortex(DVec2 center, double radius, double strength)
    {
        check(fs_add_vortex(grid_, center.x, center.y, radius, strength));
    }

    double maxDivergence() const
    {
        double v = 0.0;
        check(fs_max_divergence(grid_, &v));
        return v;
    }

    // saveCheckpoint / loadCheckpoint (FluidGridBase.h:123-125, FluidGridGPU.cpp:566-880).  A rejected
    // load throws std::runtime_error and leaves the grid unchanged.
    bool saveCheckpoint(FluidCheckpoint& checkpoint) const
    {
        size_t meta = 0, payload = 0;
        check(fs_checkpoint_size(grid_, &meta, &payload));
        std::string text(meta, '\0');
        std::vector<std::byte> data(payload);
        check(fs_checkpoint_save(grid_, text.data(), text.size(), data.data(), data.size()));
        text.resize(meta - 1);
        checkpoint.metadata = std::move(text);
        checkpoint.data = std::move(data);
        return true;
    }

    bool loadCheckpoint(const FluidCheckpoint& checkpoint)
    {
        check(fs_checkpoint_load(grid_, checkpoint.metadata.data(), checkpoint.metadata.size(),
            checkpoint.data.data(), checkpoint.data.size()));
        // refresh the adapter's view of the restored mixtures (substance ids are kept by position)
        int nl = 0, ng = 0;
        check(fs_component_counts(grid_, &nl, &ng));
        auto restore = [&](FluidMixture& mixture, int phase, int count) {
            FluidMixture out;
            for (int k = 0; k < count; ++k) {
                MixtureComponent c;
                c.substance = (size_t)k < mixture.components.size() ? mixture.components[k].substance : 0;
                check(fs_get_component(grid_, phase, k, &c.density, &c.dynamic_viscosity));
                out.components.push_back(c);
            }
            mixture = std::move(out);
        };
        restore(liquid_, 0, nl);
        restore(gas_, 1, ng);
        check(fs_get_surface_tension(grid_, &surface_tension_));
        return true;
    }

    // One FluidScene frame (FluidScene.inl:328-410): equal steps at the scene's rate, or one step when
    // the capillary subdivision reaches it; unbounded pressure budget; stops at a step not accepted.
    FrameStatistics advanceFrame(double frame_time = 0.02, int steps_per_second = 60,
        double projection_tolerance = 0.1, bool profile_stages = false)
    {
        fs_frame_settings settings{frame_time, steps_per_second, projection_tolerance, profile_stages ? 1 : 0};
        fs_frame_statistics st{};
        const fs_status status = fs_advance_frame(grid_, &settings, &st);
        FrameStatistics out;
        out.fluid_ms = st.fluid_ms;
        out.stages = {st.stages.transport_ms, st.stages.viscosity_ms, st.stages.forces_ms, st.stages.pressure_ms,
            st.stages.selection_ms, st.stages.regrid_ms};
        out.substeps = st.substeps;
        out.transport_steps = st.transport_steps;
        out.pressure_iterations = st.pressure_iterations;
        out.viscosity_iterations = st.viscosity_iterations;
        out.projection_tolerance = st.projection_tolerance;
        out.max_volume_error = st.max_volume_error;
        out.rms_volume_error = st.rms_volume_error;
        out.steps_scheduled = st.steps_scheduled;
        out.steps_run = st.steps_run;
        out.accepted = st.accepted != 0;
        check(status);
        return out;
    }

    fs_grid* handle() const { return grid_; }

    static ProjectionResult convert(const fs_projection_result& r)
    {
        ProjectionResult p;
        p.iterations = r.iterations;
        p.max_divergence = r.max_divergence;
        p.converged = r.converged != 0;
        p.physical_steps = r.physical_steps;
        p.transport_steps = r.transport_steps;
        p.relaxed = r.relaxed != 0;
        p.max_volume_error = r.max_volume_error;
        p.rms_volume_error = r.rms_volume_error;
        p.viscosity_iterations = r.viscosity_iterations;
        p.precision_limited = r.precision_limited != 0;
        return p;
    }

private:
    fs_grid* grid_ = nullptr;
    double surface_tension_ = 0.0;
    int width_;
    int height_;
    double cell_size_;
    SubstanceCatalogue substances_;
    FluidMixture liquid_, gas_;
};

// FluidGridEdit<Grid> (FluidSim/FluidGridEdit.h:97-149): a session of staged edits; only commit() touches
// the grid, and a successful commit clears the session.
template <class Grid> class FluidGridEdit {
public:
    explicit FluidGridEdit(Grid& grid) : grid_(&grid) {}
    void reserve(std::size_t count) { batch_.changes.reserve(count); }
    std::size_t size() const { return batch_.changes.size(); }
    bool empty() const { return batch_.changes.empty(); }
    void clear()
    {
        batch_.changes.clear();
        batch_.cells.clear();
        batch_.has_velocity = false;
    }
    void setCell(int x, int y, Phase phase, SubstanceId substance) { batch_.changes.push_back({x, y, phase, substance}); }
    void commit()
    {
        if (empty())
            return;
        batch_.has_velocity = false;
        grid_->commitEdits(batch_);
        clear();
    }
    void commit(DVec2 velocity)
    {
        if (empty())
            return;
        batch_.has_velocity = true;
        batch_.velocity = velocity;
        grid_->commitEdits(batch_);
        clear();
    }

private:
    Grid* grid_;
    CellEditBatch batch_;
};

inline FluidGridEdit<FluidGridHIP> FluidGridHIP::edit() { return FluidGridEdit<FluidGridHIP>(*this); }

// Batched projection (fs_batch_project; no reference counterpart): the grids share one context, each
// gets the result its own project() contract defines; one divergence tolerance per grid.
inline std::vector<FluidProjectionResult> batchProject(const std::vector<FluidGridHIP*>& grids, double dt,
    int max_iterations, const std::vector<double>& divergence_tolerances)
{
    if (divergence_tolerances.size() != grids.size())
        throw std::invalid_argument("one divergence tolerance per grid");
    std::vector<fs_grid*> handles;
    for (const FluidGridHIP* g : grids)
        handles.push_back(g ? g->handle() : nullptr);
    std::vector<fs_projection_result> raw(grids.size());
    check(fs_batch_project((int)grids.size(), handles.data(), dt, max_iterations, divergence_tolerances.data(),
        raw.data()));
    std::vector<FluidProjectionResult> out;
    for (const fs_projection_result& r : raw)
        out.push_back(FluidGridHIP::convert(r));
    return out;
}

// Red-black sweep counts for batched projections (fs_tune_batch_smoother; no reference counterpart): sets
// every grid to the fastest converging (sweeps, coarse sweeps); returns them with the time per batch.
inline std::tuple<int, int, double> tuneBatchSmoother(const std::vector<FluidGridHIP*>& grids, double dt,
    int max_iterations, const std::vector<double>& divergence_tolerances)
{
    if (divergence_tolerances.size() != grids.size())
        throw std::invalid_argument("one divergence tolerance per grid");
    std::vector<fs_grid*> handles;
    for (const FluidGridHIP* g : grids)
        handles.push_back(g ? g->handle() : nullptr);
    int sweeps = 0, coarse = 0;
    double ms = 0.0;
    check(fs_tune_batch_smoother((int)grids.size(), handles.data(), dt, max_iterations,
        divergence_tolerances.data(), &sweeps, &coarse, &ms));
    return {sweeps, coarse, ms};
}

// Batched implicit viscosity (fs_batch_apply_viscosity; no reference counterpart): each grid's
// applyViscosity(dt); returns the iteration counts.
inline std::vector<int> batchApplyViscosity(const std::vector<FluidGridHIP*>& grids, double dt)
{
    std::vector<fs_grid*> handles;
    for (const FluidGridHIP* g : grids)
        handles.push_back(g ? g->handle() : nullptr);
    std::vector<int> iterations(grids.size(), 0);
    check(fs_batch_apply_viscosity((int)grids.size(), handles.data(), dt, iterations.data()));
    return iterations;
}

// Batched timestep (fs_batch_step; no reference counterpart): each grid's step(dt[i], max_iterations[i],
// divergence_tolerance[i]) with the viscosity solves and projections batched.  A failing grid throws
// std::runtime_error after every other grid has been stepped.
inline std::vector<FluidProjectionResult> batchStep(const std::vector<FluidGridHIP*>& grids,
    const std::vector<double>& dts, const std::vector<int>& max_iterations,
    const std::vector<double>& divergence_tolerances)
{
    if (dts.size() != grids.size() || max_iterations.size() != grids.size() ||
        divergence_tolerances.size() != grids.size())
        throw std::invalid_argument("one dt, iteration budget and divergence tolerance per grid");
    std::vector<fs_grid*> handles;
    for (const FluidGridHIP* g : grids)
        handles.push_back(g ? g->handle() : nullptr);
    std::vector<fs_projection_result> raw(grids.size());
    check(fs_batch_step((int)grids.size(), handles.data(), dts.data(), max_iterations.data(),
        divergence_tolerances.data(), raw.data(), nullptr));
    std::vector<FluidProjectionResult> out;
    for (const fs_projection_result& r : raw)
        out.push_back(FluidGridHIP::convert(r));
    return out;
}

}  // namespace fluid_hip
