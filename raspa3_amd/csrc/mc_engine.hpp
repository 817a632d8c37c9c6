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
  int pad;
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
  long long trials[4];    // insertion, deletion, translation, rotation
  long long accepted[4];
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
  Proposal prop;
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
  double pTranslation, pRotation;  // swap probability = 1 - both, insertion/deletion 50/50
  double maxTranslation, maxRotation;
  long long productionCycles;
  Comp comp[MC_MAXCOMP];
};

// LDS of one wavefront in the Monte Carlo kernels: trial atom positions, Ewald eik tables and the grown molecule.
struct WaveScratch
{
  double pos[64 * 3 * (MC_MAXRIGID - 1)];
  double at[4 * 2 * MC_MAXRIGID];
  double2 tab[2 * MC_MAXRIGID * 3 * (EW_MAXK + 1)];
  double mol[3 * MC_MAXATOMS];
};

}  // namespace rg
