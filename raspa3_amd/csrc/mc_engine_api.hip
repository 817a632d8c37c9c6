// Host side of the Monte Carlo engine: multi-system GCMC state, proposal tables, hipGraph-replayed cycles, Widom.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <vector>

#include "autotune.hpp"
#include "mc_engine.hpp"

namespace rg
{
template <bool CHAIN, bool BAL, bool REINS>
__global__ void k_gcmc(const McRun* Rp, SysDev* Sarr, McSys* Yarr, int nsys, int cycles, int sample, int nw,
                       const McCtl* ctl);
__global__ void k_mc_budget(const McRun* Rp, const McSys* Yarr, int nsys, McCtl* ctl, int sample);
__global__ void k_widom_single(const McRun* Rp, const SysDev* Sp, const McSys* Yp, int comp, long long first,
                               int count, double* outLogW, double* outU, double* outEwald);
template <bool CHAIN>
__global__ void k_widom(const McRun* Rp, const SysDev* Sp, const McSys* Yp, int comp, long long first, int count,
                        double* outLogW, double* outU, double* outEwald);
__global__ void k_total_energy(const McRun* Rp, const SysDev* Sarr, const McSys* Yarr, int nsys, double* out);
__global__ void k_molecule_weights(const SysDev* Sp, Comp c, double beta, int n, const double* frac, int M,
                                   unsigned long long seed, double* logW);
}  // namespace rg

using namespace rg;

struct rg_mc
{
  rg_ctx* ctx = nullptr;
  int device = 0;
  std::vector<rg_system*> systems;
  McRun run{};
  McRun* dRun = nullptr;  // device copy read by the kernels
  SysDev* dS = nullptr;  // contiguous copies of the systems' device views (the engine's working state)
  McSys* dY = nullptr;
  std::vector<McSys> hY;  // host mirror of the static parts (pointers)
  std::vector<void*> owned;
  std::vector<std::array<void*, 3>> proposalTables;  // per system: alias probabilities, aliases, log(nvox p)
  hipGraph_t graph = nullptr;
  hipGraphExec_t graphExec = nullptr;
  int graphSample = -1;
  int graphShape = -1;  // launch-shape candidate the graph was captured with
  int forcedShape = -1;  // rg_mc_set_launch_shape: fixed candidate, or -1 for the autotuner
  int computeUnits = 256;
  McCtl* dCtl = nullptr;  // balanced launches: budget and sampled-launch index (k_mc_budget)
  bool balance = false;   // rg_mc_set_balance
  int graphBalance = -1;
};

namespace
{

rg_status buildRun(rg_ctx* ctx, const rg_forcefield* ff, int ncomp, const rg_component_desc* comps,
                   const rg_mc_params* p, McRun& R)
{
  if (ncomp < 1 || ncomp > MC_MAXCOMP) return fail(ctx, RG_ERR_UNSUPPORTED, "1..%d components", MC_MAXCOMP);
  // adsorbate atom types are staged in LDS as bytes
  if (ff->n > 255) return fail(ctx, RG_ERR_UNSUPPORTED, "the Monte Carlo engine takes at most 255 pseudo-atom types");
  const Units& u = units();
  std::memset(&R, 0, sizeof(R));
  R.seed = p->seed;
  R.beta = 1.0 / (u.kB * p->temperature);
  R.logMinW = std::log(1e-150);  // minimumRosenbluthFactor (src/raspakit/forcefield/forcefield.ixx:143)
  R.kFirst = p->k_first;
  R.kOrient = p->k_orient;
  if (R.kFirst < 1 || R.kFirst > MC_MAXTRIALS || R.kOrient < 1 || R.kOrient > MC_MAXTRIALS)
    return fail(ctx, RG_ERR_INVALID, "trial counts must lie in [1, %d]", MC_MAXTRIALS);
  R.ncomp = ncomp;
  R.maxMolecules = p->max_molecules;
  R.minStepsPerCycle = std::max(1, p->min_steps_per_cycle);
  R.pTranslation = p->p_translation;
  R.pRotation = p->p_rotation;
  R.pReinsertion = p->p_reinsertion;
  if (!(R.pTranslation >= 0.0 && R.pRotation >= 0.0 && R.pReinsertion >= 0.0 &&
        R.pTranslation + R.pRotation + R.pReinsertion <= 1.0))
    return fail(ctx, RG_ERR_INVALID, "move probabilities must be >= 0 and sum to at most 1");
  R.maxTranslation = p->max_translation;
  R.maxRotation = p->max_rotation;
  R.productionCycles = std::max(1LL, p->production_cycles);
  // pressure in internal units: Pa / (mass unit / (length unit * time unit^2)) (src/raspakit/units.cpp:54-56)
  const double pressureConversion = 1.6605402e-27 / (1e-10 * 1e-12 * 1e-12);
  double ysum = 0.0;
  for (int c = 0; c < ncomp; ++c) ysum += comps[c].mol_fraction;
  if (!(ysum > 0.0)) return fail(ctx, RG_ERR_INVALID, "mole fractions must sum to a positive value");
  for (int c = 0; c < ncomp; ++c)
  {
    const rg_component_desc& d = comps[c];
    Comp& C = R.comp[c];
    const int maxAtoms = d.flexible ? MC_MAXATOMS : MC_MAXRIGID;
    if (d.n_atoms < 1 || d.n_atoms > maxAtoms || d.first_bead < 0 || d.first_bead >= d.n_atoms)
      return fail(ctx, RG_ERR_UNSUPPORTED, "component %d: 1..%d atoms", c, maxAtoms);
    C.na = d.n_atoms;
    C.first = d.first_bead;
    C.qsum = 0.0;
    for (int a = 0; a < d.n_atoms; ++a)
    {
      if (d.type[a] < 0 || d.type[a] >= ff->n) return fail(ctx, RG_ERR_INVALID, "component %d type out of range", c);
      for (int k = 0; k < 3; ++k) C.ref[a][k] = d.position[3 * a + k] - d.position[3 * d.first_bead + k];
      C.type[a] = d.type[a];
      C.q[a] = d.charge[a];
      C.qsum += d.charge[a];
    }
    C.charged = 0;
    for (int a = 0; a < d.n_atoms; ++a) C.charged |= d.charge[a] != 0.0;
    // Ewald self and intra-molecular exclusion of one rigid molecule (src/raspakit/interactions/ewald.cpp:1165-1234)
    double self = 0.0, excl = 0.0;
    const double ps = u.coulomb * ff->alpha / std::sqrt(M_PI);
    for (int a = 0; a < d.n_atoms; ++a)
    {
      self += ps * d.charge[a] * d.charge[a];
      for (int b = a + 1; b < d.n_atoms; ++b)
      {
        double r2 = 0.0;
        for (int k = 0; k < 3; ++k) r2 += (C.ref[a][k] - C.ref[b][k]) * (C.ref[a][k] - C.ref[b][k]);
        const double r = std::sqrt(r2);
        if (r > 0.0) excl += u.coulomb * d.charge[a] * d.charge[b] * std::erf(ff->alpha * r) / r;
      }
    }
    C.eIntra = ff->use_charge ? self + excl : 0.0;
    C.index = c;
    C.flexible = d.flexible ? 1 : 0;
    if (C.flexible)
    {
      if (d.first_bead != 0 || C.charged || !(d.bond_length > 0.0) || d.bend_k_K < 0.0)
        return fail(ctx, RG_ERR_UNSUPPORTED,
                    "component %d: flexible chains grow from bead 0, are uncharged and need a positive bond length", c);
      C.bond = d.bond_length;
      C.theta0 = d.bend_theta0_deg * M_PI / 180.0;
      C.kbend = d.bend_k_K * u.kelvinToEnergy;
      for (int k = 0; k < 4; ++k) C.tors[k] = d.torsion_K[k] * u.kelvinToEnergy;
      C.logWIG = d.log_ideal_rosenbluth;
      // minimum of the torsion series on a fine grid: the rejection sampler accepts with exp(-beta (U - U_min))
      C.torsMin = INFINITY;
      for (int k = 0; k <= 7200; ++k)
      {
        const double ph = -M_PI + 2.0 * M_PI * k / 7200.0;
        const double e = C.tors[0] + C.tors[1] * (1.0 + std::cos(ph)) + C.tors[2] * (1.0 - std::cos(2.0 * ph)) +
                         C.tors[3] * (1.0 + std::cos(3.0 * ph));
        C.torsMin = std::min(C.torsMin, e);
      }
    }
    C.molFraction = d.mol_fraction / ysum;
    C.fugacity = C.molFraction * p->pressure_pa / pressureConversion;  // fugacity coefficient 1
    R.slots = std::max(R.slots, C.na);
  }
  return RG_OK;
}

}  // namespace

extern "C"
{

rg_status rg_mc_create(rg_ctx* ctx, int n_systems, rg_system* const* systems, int n_components,
                       const rg_component_desc* components, const rg_mc_params* params, rg_mc** out)
{
  if (!ctx || n_systems <= 0 || !systems || !components || !params || !out)
    return fail(ctx, RG_ERR_INVALID, "rg_mc_create: invalid arguments");
  const rg_forcefield* ff = systems[0]->ff;
  int slots = 1;
  for (int c = 0; c < n_components; ++c) slots = std::max(slots, components[c].n_atoms);
  for (int s = 0; s < n_systems; ++s)
  {
    if (!systems[s]) return fail(ctx, RG_ERR_INVALID, "rg_mc_create: system %d is null", s);
    if (systems[s]->ff != ff) return fail(ctx, RG_ERR_INVALID, "rg_mc_create: systems must share one force field");
    if (systems[s]->h.nads != 0) return fail(ctx, RG_ERR_INVALID, "rg_mc_create: systems must start empty");
    if (systems[s]->h.cap < params->max_molecules * slots)
      return fail(ctx, RG_ERR_INVALID, "rg_mc_create: system %d capacity below %d atoms", s,
                  params->max_molecules * slots);
  }
  RG_ENTER(ctx);
  rg_mc* mc = new rg_mc();
  mc->ctx = ctx;
  mc->device = ctx->device;
  if (hipDeviceGetAttribute(&mc->computeUnits, hipDeviceAttributeMultiprocessorCount, ctx->device) != hipSuccess ||
      mc->computeUnits <= 0)
    mc->computeUnits = 256;
  mc->systems.assign(systems, systems + n_systems);
  rg_status st = buildRun(ctx, ff, n_components, components, params, mc->run);
  if (st != RG_OK)
  {
    delete mc;
    return st;
  }
  auto alloc = [&](void** p, size_t bytes) -> rg_status
  {
    RG_HIP(ctx, hipMalloc(p, bytes));
    mc->owned.push_back(*p);
    return RG_OK;
  };
  std::vector<SysDev> hS(n_systems);
  mc->hY.assign(n_systems, McSys{});
  for (int s = 0; s < n_systems; ++s)
  {
    hS[s] = systems[s]->h;
    McSys& Y = mc->hY[s];
    std::memset(&Y, 0, sizeof(Y));
    void* p = nullptr;
    if ((st = alloc(&p, sizeof(int) * std::max(1, params->max_molecules))) != RG_OK) break;
    Y.comp = static_cast<int*>(p);
  }
  if (st == RG_OK) st = alloc(reinterpret_cast<void**>(&mc->dRun), sizeof(McRun));
  if (st == RG_OK) st = alloc(reinterpret_cast<void**>(&mc->dS), sizeof(SysDev) * n_systems);
  if (st == RG_OK) st = alloc(reinterpret_cast<void**>(&mc->dY), sizeof(McSys) * n_systems);
  if (st == RG_OK) st = alloc(reinterpret_cast<void**>(&mc->dCtl), sizeof(McCtl));
  if (st != RG_OK)
  {
    rg_mc_destroy(mc);
    return st;
  }
  const McCtl ctl0{};
  RG_HIP(ctx, hipMemcpyAsync(mc->dCtl, &ctl0, sizeof(McCtl), hipMemcpyHostToDevice, ctx->stream));
  RG_HIP(ctx, hipMemcpyAsync(mc->dRun, &mc->run, sizeof(McRun), hipMemcpyHostToDevice, ctx->stream));
  RG_HIP(ctx, hipMemcpyAsync(mc->dS, hS.data(), sizeof(SysDev) * n_systems, hipMemcpyHostToDevice, ctx->stream));
  RG_HIP(ctx, hipMemcpyAsync(mc->dY, mc->hY.data(), sizeof(McSys) * n_systems, hipMemcpyHostToDevice, ctx->stream));
  RG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  *out = mc;
  return RG_OK;
}

rg_status rg_mc_destroy(rg_mc* mc)
{
  if (!mc) return RG_ERR_INVALID;
  rg_status st = rgReleaseStatus(hipSetDevice(mc->device), RG_OK);
  // the context (and its stream) may already be gone: wait for the whole device instead
  st = rgReleaseStatus(hipDeviceSynchronize(), st);
  // the systems already hold the final adsorbate count and net charge (handed back after every run), so they may
  // have been released before the engine
  if (mc->graphExec) st = rgReleaseStatus(hipGraphExecDestroy(mc->graphExec), st);
  if (mc->graph) st = rgReleaseStatus(hipGraphDestroy(mc->graph), st);
  for (void* p : mc->owned) st = rgReleaseStatus(hipFree(p), st);
  for (const auto& t : mc->proposalTables)
    for (void* p : t)
      if (p) st = rgReleaseStatus(hipFree(p), st);
  delete mc;
  return st;
}

}  // extern "C"

namespace
{

// Installs the tabulated first-bead proposal of one component of one system (probability NULL: uniform again).  The
// tables it replaces are released once the device descriptor points at the new ones.
rg_status setProposal(rg_ctx* ctx, rg_mc* mc, int system, int comp, const int* n3, const double* probability,
                      const char* who)
{
  McSys& Y = mc->hY[system];
  const size_t slot = size_t(system) * MC_MAXCOMP + comp;
  if (mc->proposalTables.size() < mc->systems.size() * MC_MAXCOMP)
    mc->proposalTables.resize(mc->systems.size() * MC_MAXCOMP, {{nullptr, nullptr, nullptr}});
  const std::array<void*, 3> replaced = mc->proposalTables[slot];
  std::array<void*, 3> fresh = {{nullptr, nullptr, nullptr}};
  Proposal P{};
  if (probability)
  {
    const int nvox = n3[0] * n3[1] * n3[2];
    if (nvox <= 0 || n3[0] <= 0 || n3[1] <= 0 || n3[2] <= 0) return fail(ctx, RG_ERR_INVALID, "%s: bad voxel grid", who);
    double sum = 0.0;
    for (int v = 0; v < nvox; ++v)
    {
      if (!(probability[v] > 0.0)) return fail(ctx, RG_ERR_INVALID, "%s: probabilities must be > 0", who);
      sum += probability[v];
    }
    // Walker / Vose alias table of the normalised probabilities
    std::vector<double> prob(nvox), scaled(nvox), logp(nvox);
    std::vector<int> alias(nvox, 0);
    for (int v = 0; v < nvox; ++v)
    {
      const double pv = probability[v] / sum;
      scaled[v] = pv * nvox;
      logp[v] = std::log(pv * nvox);
    }
    std::vector<int> small, large;
    for (int v = 0; v < nvox; ++v) (scaled[v] < 1.0 ? small : large).push_back(v);
    while (!small.empty() && !large.empty())
    {
      const int s = small.back(), l = large.back();
      small.pop_back();
      prob[s] = scaled[s];
      alias[s] = l;
      scaled[l] = (scaled[l] + scaled[s]) - 1.0;
      if (scaled[l] < 1.0)
      {
        large.pop_back();
        small.push_back(l);
      }
    }
    for (int v : large) prob[v] = 1.0, alias[v] = v;
    for (int v : small) prob[v] = 1.0, alias[v] = v;
    const size_t bytes[3] = {nvox * sizeof(double), nvox * sizeof(int), nvox * sizeof(double)};
    for (int t = 0; t < 3; ++t)
    {
      const hipError_t e = hipMalloc(&fresh[t], bytes[t]);
      if (e != hipSuccess)
      {
        fresh[t] = nullptr;
        for (void* q : fresh)
          if (q) (void)hipFree(q);
        return hipfail(ctx, e, who);
      }
    }
    mc->proposalTables[slot] = fresh;
    RG_HIP(ctx, hipMemcpyAsync(fresh[0], prob.data(), bytes[0], hipMemcpyHostToDevice, ctx->stream));
    RG_HIP(ctx, hipMemcpyAsync(fresh[1], alias.data(), bytes[1], hipMemcpyHostToDevice, ctx->stream));
    RG_HIP(ctx, hipMemcpyAsync(fresh[2], logp.data(), bytes[2], hipMemcpyHostToDevice, ctx->stream));
    P.nvox = nvox;
    for (int a = 0; a < 3; ++a) P.n[a] = n3[a];
    P.aliasProb = static_cast<const double*>(fresh[0]);
    P.alias = static_cast<const int*>(fresh[1]);
    P.logNvoxP = static_cast<const double*>(fresh[2]);
  }
  else
    mc->proposalTables[slot] = fresh;
  Y.prop[comp] = P;
  // only the proposal descriptor changes: copy it into the device state
  RG_HIP(ctx, hipMemcpyAsync(reinterpret_cast<char*>(mc->dY + system) + offsetof(McSys, prop) + comp * sizeof(Proposal),
                             &Y.prop[comp], sizeof(Proposal), hipMemcpyHostToDevice, ctx->stream));
  RG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  // no kernel reads the replaced tables any more: the device descriptor points at the new ones
  for (void* q : replaced)
    if (q) RG_HIP(ctx, hipFree(q));
  return RG_OK;
}

}  // namespace

extern "C"
{

rg_status rg_mc_set_proposal(rg_ctx* ctx, rg_mc* mc, int system, const int* n3, const double* probability)
{
  if (!ctx || !mc || system < 0 || system >= int(mc->systems.size()) || (probability && !n3))
    return fail(ctx, RG_ERR_INVALID, "rg_mc_set_proposal: invalid arguments");
  RG_ENTER(ctx);
  for (int c = 0; c < mc->run.ncomp; ++c)
  {
    const rg_status st = setProposal(ctx, mc, system, c, n3, probability, "rg_mc_set_proposal");
    if (st != RG_OK) return st;
  }
  return RG_OK;
}

rg_status rg_mc_set_component_proposal(rg_ctx* ctx, rg_mc* mc, int system, int component, const int* n3,
                                       const double* probability)
{
  if (!ctx || !mc || system < 0 || system >= int(mc->systems.size()) || component < 0 ||
      component >= mc->run.ncomp || (probability && !n3))
    return fail(ctx, RG_ERR_INVALID, "rg_mc_set_component_proposal: invalid arguments");
  RG_ENTER(ctx);
  return setProposal(ctx, mc, system, component, n3, probability, "rg_mc_set_component_proposal");
}

rg_status rg_mc_set_launch_shape(rg_ctx* ctx, rg_mc* mc, int shape)
{
  if (!ctx || !mc || shape < -1 || shape > 3) return fail(ctx, RG_ERR_INVALID, "rg_mc_set_launch_shape: shape in [-1, 3]");
  mc->forcedShape = shape;
  return RG_OK;
}

rg_status rg_mc_set_balance(rg_ctx* ctx, rg_mc* mc, int enable, double overhead_molecules)
{
  if (!ctx || !mc || !std::isfinite(overhead_molecules) || overhead_molecules < 0.0)
    return fail(ctx, RG_ERR_INVALID, "rg_mc_set_balance: invalid arguments");
  RG_ENTER(ctx);
  for (int c = 0; enable && c < mc->run.ncomp; ++c)
    if (mc->run.comp[c].flexible)
      return fail(ctx, RG_ERR_UNSUPPORTED, "rg_mc_set_balance: balanced launches are built for rigid molecules only");
  mc->balance = enable != 0;
  mc->run.balanceOverhead = overhead_molecules;
  RG_HIP(ctx, hipMemcpyAsync(mc->dRun, &mc->run, sizeof(McRun), hipMemcpyHostToDevice, ctx->stream));
  RG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return RG_OK;
}

rg_status rg_mc_run(rg_ctx* ctx, rg_mc* mc, int cycles, int sample, int use_graph)
{
  if (!ctx || !mc || cycles < 0) return fail(ctx, RG_ERR_INVALID, "rg_mc_run: invalid arguments");
  RG_ENTER(ctx);
  const int nsys = static_cast<int>(mc->systems.size());
  // The launch shape is tuned online: one wavefront per system with 1, 2 or 4 systems per workgroup, or a team of
  // MC_PARTS wavefronts per system (one system per workgroup).  While the key is open, cycles of the real run are
  // launched eagerly with each candidate and timed.  Every candidate computes the same trajectory bit for bit (the
  // heavy loops are cut into MC_PARTS fixed partitions combined in a fixed order), so the choice only affects speed.
  static const int kNw[4] = {1, 1, 1, MC_PARTS};
  static const int kWpb[4] = {1, 2, 4, 1};
  bool chain = false;
  for (int c = 0; c < mc->run.ncomp; ++c) chain = chain || mc->run.comp[c].flexible;
  const bool bal = mc->balance && !chain;
  auto launch = [&](int cand)
  {
    const int nw = kNw[cand], wpb = kWpb[cand];
    if (bal) hipLaunchKernelGGL(k_mc_budget, dim3(1), dim3(256), 0, ctx->stream, mc->dRun, mc->dY, nsys, mc->dCtl, sample);
    // kernel instances: chains (no balancing) or rigid molecules, each with or without the reinsertion move
    const bool reins = mc->run.pReinsertion > 0.0;
    auto kernel = chain  ? (reins ? k_gcmc<true, false, true> : k_gcmc<true, false, false>)
                  : bal  ? (reins ? k_gcmc<false, true, true> : k_gcmc<false, true, false>)
                         : (reins ? k_gcmc<false, false, true> : k_gcmc<false, false, false>);
    hipLaunchKernelGGL(kernel, dim3((nsys + wpb - 1) / wpb), dim3(64 * nw * wpb), nw * wpb * sizeof(WaveScratch),
                       ctx->stream, mc->dRun, mc->dS, mc->dY, nsys, 1, sample, nw, static_cast<const McCtl*>(mc->dCtl));
    return hipGetLastError();
  };
  Autotuner& tuner = Autotuner::instance();
  // the best shape depends on the loading (a team's barriers only pay off once the systems hold molecules), so the
  // key carries the total adsorbate atom count rounded down to a power of two (host copy of the last run's count)
  long long atoms = 0;
  for (const rg_system* sy : mc->systems) atoms += sy->h.nads;
  int bucket = 0;
  while ((2ll << bucket) <= atoms) ++bucket;
  const std::string key = "k_gcmc/systems=" + std::to_string(nsys) + "/cu=" + std::to_string(mc->computeUnits) +
                          "/atoms>=" + std::to_string(atoms > 0 ? (1ll << bucket) : 0) + (bal ? "/balanced" : "");
  int done = 0;
  for (int cand; mc->forcedShape < 0 && done < cycles && (cand = tuner.next(key, 4, 2)) >= 0; ++done)
  {
    float ms = 0.0f;
    RG_HIP(ctx, timeLaunch(ctx->stream, [&] { return launch(cand); }, ms));
    tuner.report(key, cand, ms);
  }
  const int best = tuner.best(key);
  const int shape = mc->forcedShape >= 0 ? mc->forcedShape : (best >= 0 ? best : 0);
  if (!use_graph)
  {
    for (int c = done; c < cycles; ++c) RG_HIP(ctx, launch(shape));
  }
  else if (done < cycles)
  {
    // one MC cycle of every system (with its budget kernel when balanced) captured once as a graph, replayed per
    // cycle
    if (!mc->graphExec || mc->graphSample != sample || mc->graphShape != shape || mc->graphBalance != int(bal))
    {
      if (mc->graphExec) (void)hipGraphExecDestroy(mc->graphExec);
      if (mc->graph) (void)hipGraphDestroy(mc->graph);
      mc->graphExec = nullptr;
      mc->graph = nullptr;
      RG_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
      const hipError_t le = launch(shape);
      hipError_t e = hipStreamEndCapture(ctx->stream, &mc->graph);
      if (le != hipSuccess) return hipfail(ctx, le, "k_gcmc launch under capture");
      if (e != hipSuccess) return hipfail(ctx, e, "hipStreamEndCapture");
      RG_HIP(ctx, hipGraphInstantiate(&mc->graphExec, mc->graph, nullptr, nullptr, 0));
      mc->graphSample = sample;
      mc->graphShape = shape;
      mc->graphBalance = int(bal);
    }
    for (int c = done; c < cycles; ++c) RG_HIP(ctx, hipGraphLaunch(mc->graphExec, ctx->stream));
  }
  // hand the adsorbate count and net charge back to the systems (their Ewald state and adsorbate arrays are
  // updated in place by the kernel)
  std::vector<SysDev> hS(nsys);
  RG_HIP(ctx, hipMemcpyAsync(hS.data(), mc->dS, sizeof(SysDev) * nsys, hipMemcpyDeviceToHost, ctx->stream));
  RG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  for (int s = 0; s < nsys; ++s)
  {
    rg_system* sy = mc->systems[s];
    sy->h.nads = hS[s].nads;
    sy->h.netCharge = hS[s].netCharge;
    RG_HIP(ctx, hipMemcpyAsync(sy->d, &sy->h, sizeof(SysDev), hipMemcpyHostToDevice, ctx->stream));
  }
  RG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return RG_OK;
}

rg_status rg_mc_state(rg_ctx* ctx, const rg_mc* mc, int system, rg_mc_state_t* out)
{
  if (!ctx || !mc || !out || system < 0 || system >= int(mc->systems.size()))
    return fail(ctx, RG_ERR_INVALID, "rg_mc_state: invalid arguments");
  RG_ENTER(ctx);
  McSys Y;
  RG_HIP(ctx, hipMemcpyAsync(&Y, mc->dY + system, sizeof(McSys), hipMemcpyDeviceToHost, ctx->stream));
  RG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  std::memset(out, 0, sizeof(*out));
  out->n_molecules = Y.nmol;
  for (int c = 0; c < MC_MAXCOMP; ++c) out->n_per_component[c] = Y.ncomp[c];
  out->energy = Y.U;
  out->steps = Y.step;
  out->cycles_sampled = Y.cycle;
  for (int q = 0; q < 5; ++q)
  {
    out->trials[q] = Y.stat.trials[q];
    out->accepted[q] = Y.stat.accepted[q];
  }
  for (int b = 0; b < MC_BLOCKS; ++b)
  {
    out->block_count[b] = Y.block[b].count;
    out->block_energy[b] = Y.block[b].U;
    for (int c = 0; c < MC_MAXCOMP; ++c)
    {
      out->block_n[b][c] = Y.block[b].N[c];
      out->block_energy_n[b][c] = Y.block[b].UN[c];
      for (int d = 0; d < MC_MAXCOMP; ++d) out->block_nn[b][c][d] = Y.block[b].NN[c][d];
    }
  }
  return RG_OK;
}

rg_status rg_mc_adsorbates(rg_ctx* ctx, const rg_mc* mc, int system, int capacity, double* pos, int* type,
                           int* component, int* n_out)
{
  if (!ctx || !mc || !n_out || system < 0 || system >= int(mc->systems.size()))
    return fail(ctx, RG_ERR_INVALID, "rg_mc_adsorbates: invalid arguments");
  RG_ENTER(ctx);
  McSys Y;
  SysDev S;
  RG_HIP(ctx, hipMemcpyAsync(&Y, mc->dY + system, sizeof(McSys), hipMemcpyDeviceToHost, ctx->stream));
  RG_HIP(ctx, hipMemcpyAsync(&S, mc->dS + system, sizeof(SysDev), hipMemcpyDeviceToHost, ctx->stream));
  RG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  const int n = Y.nmol * mc->run.slots;
  *n_out = n;
  if (n > capacity || n == 0) return n > capacity ? fail(ctx, RG_ERR_INVALID, "capacity below %d", n) : RG_OK;
  std::vector<double> x(n), y(n), z(n);
  std::vector<int> t(n), c(Y.nmol);
  RG_HIP(ctx, hipMemcpyAsync(x.data(), S.ax, n * 8, hipMemcpyDeviceToHost, ctx->stream));
  RG_HIP(ctx, hipMemcpyAsync(y.data(), S.ay, n * 8, hipMemcpyDeviceToHost, ctx->stream));
  RG_HIP(ctx, hipMemcpyAsync(z.data(), S.az, n * 8, hipMemcpyDeviceToHost, ctx->stream));
  RG_HIP(ctx, hipMemcpyAsync(t.data(), S.at, n * 4, hipMemcpyDeviceToHost, ctx->stream));
  RG_HIP(ctx, hipMemcpyAsync(c.data(), Y.comp, Y.nmol * 4, hipMemcpyDeviceToHost, ctx->stream));
  RG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  for (int a = 0; a < n; ++a)
  {
    if (pos)
    {
      pos[3 * a] = x[a];
      pos[3 * a + 1] = y[a];
      pos[3 * a + 2] = z[a];
    }
    if (type) type[a] = t[a];
  }
  if (component)
    for (int m = 0; m < Y.nmol; ++m) component[m] = c[m];
  return RG_OK;
}

rg_status rg_mc_total_energy(rg_ctx* ctx, const rg_mc* mc, double* out)
{
  if (!ctx || !mc || !out) return fail(ctx, RG_ERR_INVALID, "rg_mc_total_energy: invalid arguments");
  RG_ENTER(ctx);
  const int nsys = static_cast<int>(mc->systems.size());
  double* d = nullptr;
  RG_HIP(ctx, hipMalloc(&d, nsys * sizeof(double)));
  hipLaunchKernelGGL(k_total_energy, dim3((nsys + 3) / 4), dim3(256), 0, ctx->stream, mc->dRun, mc->dS, mc->dY,
                     nsys, d);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(out, d, nsys * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  (void)hipStreamSynchronize(ctx->stream);
  (void)hipFree(d);
  if (e != hipSuccess) return hipfail(ctx, e, "rg_mc_total_energy");
  return RG_OK;
}

// Launches the Widom kernels into device output arrays on the context stream.  'scratch' (runBytes + sysBytes of
// device memory) receives the run and system descriptors; the caller releases it after the stream has drained.
static rg_status widomLaunch(rg_ctx* ctx, const rg_system* sys, const rg_component_desc* component,
                             const rg_mc_params* params, long long first, int count, double* dLogW, double* dU,
                             double* dEw, void** scratch)
{
  McRun R;
  rg_status st = buildRun(ctx, sys->ff, 1, component, params, R);
  if (st != RG_OK) return st;
  const size_t runBytes = (sizeof(McRun) + 255) / 256 * 256;
  RG_HIP(ctx, hipMalloc(scratch, runBytes + sizeof(McSys)));
  McSys Y;
  std::memset(&Y, 0, sizeof(Y));
  McRun* dR = static_cast<McRun*>(*scratch);
  McSys* dY = reinterpret_cast<McSys*>(static_cast<char*>(*scratch) + runBytes);
  RG_HIP(ctx, hipMemcpyAsync(dR, &R, sizeof(McRun), hipMemcpyHostToDevice, ctx->stream));
  RG_HIP(ctx, hipMemcpyAsync(dY, &Y, sizeof(McSys), hipMemcpyHostToDevice, ctx->stream));
  // a single uncharged atom on a grid: several insertions per wavefront (k_widom_single); otherwise one
  const int t0 = component->type[0];
  const bool single = component->n_atoms == 1 && !component->flexible &&
                      (component->charge[0] == 0.0 || !sys->ff->use_charge) && t0 >= 0 &&
                      t0 < int(sys->gridOfType.size()) && sys->gridOfType[t0] >= 0 && R.kFirst <= 64;
  if (single)
  {
    const int perWave = 64 / R.kFirst;
    const long long waves = (static_cast<long long>(count) + perWave - 1) / perWave;
    hipLaunchKernelGGL(k_widom_single, dim3(static_cast<unsigned>((waves + 3) / 4)), dim3(256), 0, ctx->stream, dR,
                       sys->d, dY, 0, first, count, dLogW, dU, dEw);
  }
  else
    hipLaunchKernelGGL(component->flexible ? k_widom<true> : k_widom<false>, dim3((count + 3) / 4), dim3(256), 0,
                       ctx->stream, dR, sys->d, dY, 0, first, count, dLogW, dU, dEw);
  RG_HIP(ctx, hipGetLastError());
  return RG_OK;
}

rg_status rg_widom(rg_ctx* ctx, const rg_system* sys, const rg_component_desc* component, const rg_mc_params* params,
                   long long first, int count, double* log_weight, double* energy, double* ewald)
{
  if (!ctx || !sys || !component || !params || count < 0 || (count > 0 && (!log_weight || !energy || !ewald)))
    return fail(ctx, RG_ERR_INVALID, "rg_widom: invalid arguments");
  if (count == 0) return RG_OK;
  RG_ENTER(ctx);
  double* out = nullptr;
  void* scratch = nullptr;
  RG_HIP(ctx, hipMalloc(&out, 3 * size_t(count) * sizeof(double)));
  rg_status st = widomLaunch(ctx, sys, component, params, first, count, out, out + count, out + 2 * size_t(count),
                             &scratch);
  hipError_t e = hipSuccess;
  if (st == RG_OK) e = hipMemcpyAsync(log_weight, out, count * 8, hipMemcpyDeviceToHost, ctx->stream);
  if (st == RG_OK && e == hipSuccess)
    e = hipMemcpyAsync(energy, out + count, count * 8, hipMemcpyDeviceToHost, ctx->stream);
  if (st == RG_OK && e == hipSuccess)
    e = hipMemcpyAsync(ewald, out + 2 * size_t(count), count * 8, hipMemcpyDeviceToHost, ctx->stream);
  if (st == RG_OK && e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  (void)hipStreamSynchronize(ctx->stream);
  (void)hipFree(out);
  if (scratch) (void)hipFree(scratch);
  if (st != RG_OK) return st;
  if (e != hipSuccess) return hipfail(ctx, e, "rg_widom");
  return RG_OK;
}

rg_status rg_widom_device(rg_ctx* ctx, const rg_system* sys, const rg_component_desc* component,
                          const rg_mc_params* params, long long first, int count, double* d_log_weight,
                          double* d_energy, double* d_ewald)
{
  if (!ctx || !sys || !component || !params || count < 0 ||
      (count > 0 && (!d_log_weight || !d_energy || !d_ewald)))
    return fail(ctx, RG_ERR_INVALID, "rg_widom_device: invalid arguments");
  if (count == 0) return RG_OK;
  RG_ENTER(ctx);
  void* scratch = nullptr;
  const rg_status st = widomLaunch(ctx, sys, component, params, first, count, d_log_weight, d_energy, d_ewald,
                                   &scratch);
  // the descriptors are read by the kernel: released once the stream has passed it
  if (scratch)
  {
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(scratch);
  }
  return st;
}

rg_status rg_molecule_weights(rg_ctx* ctx, const rg_system* sys, const rg_component_desc* component,
                              double temperature, int n, const double* frac, int n_orient, unsigned long long seed,
                              double* log_weight)
{
  if (!ctx || !sys || !component || n < 0 || (n > 0 && (!frac || !log_weight)) || !(temperature > 0.0) ||
      n_orient < 1 || n_orient > 64 || (n_orient & (n_orient - 1)))
    return fail(ctx, RG_ERR_INVALID, "rg_molecule_weights: invalid arguments (n_orient: a power of two up to 64)");
  if (n == 0) return RG_OK;
  rg_mc_params params;
  std::memset(&params, 0, sizeof(params));
  params.temperature = temperature;
  params.k_first = params.k_orient = 1;
  McRun R;
  rg_status st = buildRun(ctx, sys->ff, 1, component, &params, R);
  if (st != RG_OK) return st;
  for (int a = 0; a < component->n_atoms; ++a)
  {
    const int t = component->type[a];
    if (t >= int(sys->gridOfType.size()) || sys->gridOfType[t] < 0)
      return fail(ctx, RG_ERR_UNSUPPORTED, "rg_molecule_weights: atom %d (type %d) has no interpolation grid", a, t);
  }
  RG_ENTER(ctx);
  void* d = nullptr;
  RG_HIP(ctx, hipMalloc(&d, size_t(n) * 4 * sizeof(double)));
  double* dfrac = static_cast<double*>(d);
  double* dout = dfrac + 3 * size_t(n);
  hipError_t e = hipMemcpyAsync(dfrac, frac, size_t(n) * 3 * sizeof(double), hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess)
  {
    const size_t threads = size_t(n) * n_orient;
    hipLaunchKernelGGL(k_molecule_weights, dim3(unsigned((threads + 255) / 256)), dim3(256), 0, ctx->stream, sys->d,
                       R.comp[0], R.beta, n, dfrac, n_orient, seed, dout);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(log_weight, dout, size_t(n) * sizeof(double), hipMemcpyDeviceToHost,
                                          ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  (void)hipStreamSynchronize(ctx->stream);
  (void)hipFree(d);
  if (e != hipSuccess) return hipfail(ctx, e, "rg_molecule_weights");
  return RG_OK;
}

}  // extern "C"
