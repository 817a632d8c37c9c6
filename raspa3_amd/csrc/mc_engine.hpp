// Monte Carlo engine data shared between the device kernels (mc_engine.hip) and the host API (mc_engine_api.hip).
#pragma once

#include "mc_internal.hpp"

#define MC_MAXATOMS 8   // atom slots per molecule in the adsorbate arrays (flexible chains up to 8 beads)
#define MC_MAXRIGID 4   // atoms of a rigid molecule
#define MC_MAXREJ 128   // rejection attempts per sampled bend or torsion angle of a chain
#define MC_MAXCOMP 2    // components per system
#define MC_BLOCKS 5     // block averages (src/raspakit/properties/property_block_average.ixx)
#define MC_MAXTRIALS 64 // first-bead / orientation trials (one lane each)

namespace rg
{

// A rigid component: body-frame atom positions relative to the first bead, types, charges, and the constant
// Ewald self + intra-molecular exclusion energy of one molecule.
struct Comp
{
  int na;
  int first;
  int charged;
  int index;          // position of the component in the run (selects its first-bead proposal)
  double ref[MC_MAXATOMS][3];
  int type[MC_MAXATOMS];
  double q[MC_MAXATOMS];
  double qsum;
  double eIntra;      // C alpha/sqrt(pi) sum q^2 + sum_pairs C qi qj erf(alpha r)/r  (internal units)
  double molFraction;
  double fugacity;    // y phi P in internal units
  // flexible linear chain (first bead 0, grown 1, 2, ...): bond length, harmonic bend (theta0 in rad, k in internal
  // units / rad^2), torsion series c0 + c1 (1 + cos p) + c2 (1 - cos 2p) + c3 (1 + cos 3p) and its minimum, log of
  // the ideal-gas Rosenbluth weight of the growth
  int flexible;
  int pad2;
  double bond, theta0, kbend, tors[4], torsMin, logWIG;
};

// Tabulated first-bead proposal: probabilities p_v over nvox voxels of the cell, alias table, and
// log(nvox p_v) (the log of V q on voxel v).
struct Proposal
{
  int nvox;
  int n[3];
  const double* aliasProb;
  const int* alias;
  const double* logNvoxP;
};

struct McBlock
{
  double count;
  double U;
  double N[MC_MAXCOMP];
  double UN[MC_MAXCOMP];
  double NN[MC_MAXCOMP][MC_MAXCOMP];
};

struct McStat
{
  long long trials[5];    // insertion, deletion, translation, rotation, reinsertion
  long long accepted[5];
};

// Per-system mutable Monte Carlo state (device).
struct McSys
{
  int nmol;
  int ncomp[MC_MAXCOMP];
  int pad;
  double U;               // running adsorbate energy (framework-molecule + molecule-molecule + Ewald), internal units
  long long step;         // move counter = Philox stream position
  long long cycle;        // production cycles sampled
  int* comp;              // component of each molecule slot
  Proposal prop[MC_MAXCOMP];  // first-bead proposal per component (nvox = 0: uniform)
  McStat stat;
  McBlock block[MC_BLOCKS];
};

// Run parameters shared by all systems of a launch (kernel argument).
struct McRun
{
  unsigned long long seed;
  double beta;
  double logMinW;         // log(minimumRosenbluthFactor)
  int kFirst, kOrient;
  int ncomp;
  int maxMolecules;
  int minStepsPerCycle;
  int slots;              // atom slots per molecule in the adsorbate arrays: the largest n_atoms of the components
  double pTranslation, pRotation, pReinsertion;  // swap probability = 1 - all three, insertion/deletion 50/50
  double maxTranslation, maxRotation;
  long long productionCycles;
  double balanceOverhead;  // balanced launches: a cycle's cost is its steps times (molecules + this)
  Comp comp[MC_MAXCOMP];
};

// Per-launch control of balanced Monte Carlo launches, written by k_mc_budget before each k_gcmc launch: every
// system runs whole cycles while their summed cost stays within 'budget' (the largest single-cycle cost over the
// systems at the start of the launch); 'sampledIndex' is the launch's index among the sampled launches, which picks
// the block its samples go to.
struct McCtl
{
  double budget;
  int sampledIndex;
  int sampledCount;
};
// at most this many cycles of one system per balanced launch
#define MC_MAX_BALANCED_CYCLES 64

// LDS of one wavefront in the Monte Carlo kernels: trial atom positions, Ewald eik tables and the grown molecule.
// sites of one energy batch (K trial sets of up to MC_MAXRIGID - 1 sites), candidate pairs per flush, and force
// fields whose pair table is copied into LDS (larger ones read it from global memory).  The scratch is kept under
// 20 KB per wavefront so that two workgroups of four wavefronts fit the 160 KB of LDS of a compute unit.
#define MC_MAXSITES (64 * (MC_MAXRIGID - 1))
#define MC_PAIRCAP 128
#define MC_LDSTYPES 12

struct WaveScratch
{
  double pos[64 * 3 * (MC_MAXRIGID - 1)];
  double at[4 * 2 * MC_MAXRIGID];
  double2 tab[2 * MC_MAXRIGID * 3 * (EW_MAXK + 1)];
  double mol[3 * MC_MAXATOMS];
  // batched site energies: per-site charge, energy and the current partition's sum (siteT); candidate pairs
  // (r^2, key = site << 8 | lane) of the current chunk of 64 adsorbate atoms, whose charges and types are chQ / chT
  double siteE[MC_MAXSITES];
  double sQ[MC_MAXSITES];
  double siteT[MC_MAXSITES];
  double pRR[MC_PAIRCAP];
  double chQ[64];
  double deT;  // one partition's Ewald difference
  short sType[MC_MAXSITES];
  unsigned short pKey[MC_PAIRCAP];
  unsigned char siteOv[MC_MAXSITES];
  unsigned char ovT[MC_MAXSITES];
  unsigned char chT[64];
};
static_assert(sizeof(WaveScratch) <= 19 * 1024, "four WaveScratch and the LDS pair table must fit in 80 KB");

// The heavy loops of a move (adsorbate atom chunks, Ewald k-vector passes) are cut into MC_PARTS fixed partitions
// whose results are combined in partition order, so a system's trajectory is the same bit for bit whether one
// wavefront runs all partitions or a team of MC_PARTS wavefronts runs one each.
#define MC_PARTS 4

// The wavefronts that run one system: nw (1 or MC_PARTS) of them, this one is wi; ws is the team's first
// WaveScratch (the team's are consecutive).
struct Team
{
  int nw, wi;
  WaveScratch* ws;
};

}  // namespace rg
