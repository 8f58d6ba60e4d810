/*
 * host_env.h -- the state of a minimal host program for testing the
 * reference-side binding (tandav_amd/glue/tandav_glue.c) without Tandav.
 *
 * The binding is written against Tandav's globals; this test host declares
 * the ones it reads, with Tandav's names, types and layouts (globals.h:17-36,
 * particles.h:17-37, domain.h:24-63, Gravity/tree.h:23-59, signal.h:10-23,
 * macro.h:13-15), and host.c defines them and drives the binding's team
 * functions from an OpenMP team the way Tandav's main loop does.  Config
 * knobs that Tandav fixes at compile time (TREE_OPEN_PARAM_*) are
 * variables here so one build serves every test.
 */
#ifndef TDV_TEST_HOST_ENV_H
#define TDV_TEST_HOST_ENV_H

#include <stdlib.h>
#include <stdio.h>
#include <stdint.h>
#include <stdbool.h>
#include <string.h>
#include <math.h>
#include <omp.h>

#define DOUBLE_PRECISION
#define GRAVITY
#define GRAVITY_TREE
#define VECTOR_SIZE 64
#define NPARTYPE 6
#define TREE_OPEN_PARAM_BH Host_Theta_BH
#define TREE_OPEN_PARAM_REL Host_Alpha_Rel
extern double Host_Theta_BH, Host_Alpha_Rel;

typedef double Float;
typedef uint32_t ID_t;
typedef uint32_t intime_t;
typedef uint64_t shortKey;
typedef __uint128_t peanoKey;

#define N_PEANO_BITS 128
#define N_SHORT_BITS 64
#define DELTA_PEANO_BITS (N_PEANO_BITS - N_SHORT_BITS)

extern struct Particle_Data {
	int * restrict Type;
	int * restrict Time_Bin;
	intime_t * restrict It_Drift_Pos;
	intime_t * restrict It_Kick_Pos;
	peanoKey * restrict Key;
	ID_t * restrict ID;
	Float * restrict Cost;
	Float * restrict Pos[3];
	Float * restrict Vel[3];
	Float * restrict Acc[3];
	Float * restrict Mass;
	Float * restrict Grav_Acc[3];
	Float * restrict Last_Acc_Mag;
	int * restrict Tree_Parent;
} P;

extern union Domain_Node_List {
	struct Bunch_Node {
		shortKey Key;
		int Target;
		int Level;
		int First_Part;
		uint64_t Npart;
		float Cost;
		bool Is_Local;
		int Modify;
	} Bunch;
	struct Top_Tree_Node {
		shortKey Key;
		int Target;
		int Level;
		int First_Part;
		int Npart;
		float Pos[3];
		float Mass;
		float CoM[3];
		float Dp[3];
	} TNode;
} * restrict D;

extern uint32_t NTop_Nodes;

extern struct Domain_Properties {
	double Size;
	double Origin[3];
	double Center[3];
	double Center_Of_Mass[3];
} Domain;

extern struct Tree_Node {
	int DNext;
	uint32_t Bitfield;
	int DUp;
	int Npart;
	Float Pos[3];
	Float Mass;
	Float CoM[3];
	Float Dp[3];
} * restrict Tree;

extern double Epsilon[NPARTYPE], Epsilon2[NPARTYPE], Epsilon3[NPARTYPE];
extern uint32_t NNodes;

enum Tree_Bitfield { LOCAL = 9, TOP = 10, UPDATED = 11 };

extern struct Simulation_Signals {
	bool Sync_Point;
	bool Write_Snapshot;
	bool Restart_Write_File;
	bool Restart_Continue;
	bool Endrun;
	bool Prepare_Step;
	bool First_Step;
	bool Domain_Update;
	bool Tree_Update;
	bool Use_BH_Criterion;
} Sig;
#pragma omp threadprivate(Sig)

extern struct Local_Task_Properties {
	int ID;
	int Rank;
	int Thread_ID;
	bool Is_Master;
	int Npart_Total;
	int Npart[NPARTYPE];
	unsigned short Seed[3];
} Task;
#pragma omp threadprivate(Task)

extern struct Global_Simulation_Properties {
	uint64_t Npart_Total;
	double Boxsize[3];
} Sim;

extern struct Parameters_From_File {
	double Grav_Softening[NPARTYPE];
} Param;

extern struct Constants_In_Code_Units {
	double Gravity;
} Const;

extern int * restrict Active_Particle_List;
extern int NActive_Particles;

void Make_Active_Particle_List(void);
double erand48(unsigned short xsubi[3]);

/* the binding's exports, with the signatures Tandav declares */
void Compute_Acceleration(void);
void Sort_Particles_By_Peano_Key(void);
void Reverse_Peano_Keys(void);
int cmp_peanoKeys(const void *a, const void *b);
peanoKey Peano_Key(const Float px, const Float py, const Float pz);
peanoKey Reversed_Peano_Key(const Float px, const Float py, const Float pz);
peanoKey Reverse_Peano_Key(const peanoKey pkey);
shortKey Short_Peano_Key(const Float px, const Float py, const Float pz);
shortKey Reversed_Short_Peano_Key(const Float px, const Float py,
		const Float pz);
void Test_Peanokey(void);
void Setup_Gravity_Tree(void);
void Gravity_Tree_Build(void);
void Gravity_Tree_Acceleration(void);
void Gravity_Tree_Periodic(void);
void Gravity_Tree_Update_Kicks(const int ipart, const double dt);
void Gravity_Tree_Update_Drift(const double dt);
void Gravity_Tree_Free(void);
Float Node_Size(const int node);
int Level(const int node);
bool Node_Is(const enum Tree_Bitfield bit, const int node);
void Node_Set(const enum Tree_Bitfield bit, const int node);
void Node_Clear(const enum Tree_Bitfield bit, const int node);
#ifdef PERIODIC
void Gravity_Periodic_Init(void);
void Ewald_Correction(const Float dr[3], Float f[3]);
#endif
#ifdef GRAVITY_FORCETEST
void Gravity_Forcetest(void);
#endif

#define rprintf(...) do { if (Task.Is_Master) printf(__VA_ARGS__); } while (0)
#define Assert(expr, ...) do { if (!(expr)) { \
	fprintf(stderr, "%s:%d: ", __FILE__, __LINE__); \
	fprintf(stderr, __VA_ARGS__); fputc('\n', stderr); abort(); } } while (0)

#endif
