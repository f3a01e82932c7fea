// The reference's initial-data plugins (src/Cell/initial_data/*.c) on the
// host: the global functions that set a disk's initial primitives, and the
// per-annulus rows the fixed r boundaries and the boundary damping read on
// the device.
#pragma once
#include "disco_internal.h"

// The global function of initial-data type `type` (cell_init_ptr,
// src/Cell/cell_init_function_ptr.c:6-25) for one disk: the primitives
// (q-major [NQ][nz], the reference's zone order) from the cells' tiph / dphi
// and the masses g (null: as gravMass_init_ptr sets them, which src/main.c:103
// does before the cells).  Zones and quantities the plugin leaves alone keep
// their values.  Returns false with the error set for an unknown type.
bool disco_cell_init(const Sim *S, int type, const GravMass *g, const double *tiph, const double *dphi, double *prim,
                     long long nz, int NQ);

// The single-cell function of type `type` (cell_single_init_ptr, :27-46) for
// one cell of annulus i with edge tiph and width dphi: writes what the plugin
// sets into prim[0..NQ) and returns how many of the first five it set (0 or
// 5); the passive slot NUM_C is written where the plugin writes it.  -1 with
// the error set for an unknown type and for the torus, whose single-cell
// function exits.
int disco_single_init(const Sim *S, int type, int i, double tiph, double dphi, double *prim, int NQ);

// Rows of every annulus (k-major, as the annulus table) of one disk for the
// fixed r boundaries and the damping; false with the error set for the
// combinations the reference cannot run (the torus with a fixed boundary or
// damping; damping whose single-cell values depend on the uninitialised
// cell cell_bc_damp passes, src/Cell/cell_misc.c:107,123).  *any is set when
// some row asks for device work.
// values_always: the rows' single-cell values also where the disk's set-up
// uses no fixed boundary or damping (the class API's unit-level calls).
bool disco_init_rows(const Sim *S, int NQ, std::vector<InitRow> &rows, bool *any, bool values_always = false);
