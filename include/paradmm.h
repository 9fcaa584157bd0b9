/*
 * paradmm.h -- C API of the parADMM engine (MI355X build).
 *
 * Source-compatible with the reference header core/paradmm.h: the same
 * `graph` struct (core/paradmm.h:33-63), field for field, and the same 28
 * prototypes (core/paradmm.h:69-115), so that a caller written against the
 * reference (packing/host.c) compiles and links against libparadmm unchanged.
 *
 * What differs behind the API: the update passes run on the GPU (HIP kernels
 * over structure-of-arrays fp32 edge buffers).  Each legacy update call keeps
 * the host arrays coherent: it uploads what the pass reads and downloads what
 * it writes, so a caller may read or write G.X, G.Z, G.RHOS, ... between calls
 * exactly as with the reference.  The extension header paradmm_gpu.h adds
 * device-resident iteration, residuals, batched solves and learned rho.
 *
 * Prox callbacks keep the reference signature
 *   void op(double *x, double *n, double *rhos, int nvars, int ndims, void *params)
 * but a GPU cannot call a host function: each callback address must be a
 * known device prox kind (the operators exported by paradmm_gpu.h, the
 * reference demo's operators identified by symbol name, or any callback
 * registered with paradmm_register_prox_kind).
 */
#ifndef __PARADMM_H
#define __PARADMM_H

#ifdef NEEDS_EXTERN_C
#define MY_EXTERN_C extern "C"
#else
#define MY_EXTERN_C
#endif

/* The reference build passes -DSILENCE=0|1 (SConstruct:23); callers such as
 * packing/host.c test it.  Default to quiet when the build does not set it. */
#ifndef SILENCE
#define SILENCE 1
#endif

#include <stddef.h>

/* unused by the engine; kept for source compatibility (core/paradmm.h:22-31) */
typedef struct params_matrix
{
	int flag_1stiter;
	int rows;
	double* A;
	double* B;
	double* R;
	double* ATB;
	double* ATA;
} mparams;

/* core/paradmm.h:33-63 */
typedef struct graphs
{
	int num_opts;   /* number of factors (operators) */
	int num_vars;   /* number of variables */
	int num_dims;   /* dimension of every variable */

	int num_edges;  /* number of edges */

	void (**list_of_operators_functions_pointers)(double*, double*, double*, int, int, void*);
	void** list_of_operators_param_pointers;
	int* list_of_operators_param_sizes;

	int* operatorix_to_numedges;
	int** operatorix_to_edgelistix;

	int* varix_to_numedges;
	int** varix_to_edgelistix;

	int* edgeix_to_varlistix;

	double* X;  /* num_dims * num_edges, edge-major (X[num_dims*e + k]) */
	double* M;
	double* N;
	double* U;

	double* Z;  /* num_dims * num_vars */

	double* RHOS;    /* num_edges */
	double* ALPHAS;  /* num_edges */

} graph;

/* utilities (core/paradmm.h:69-74) */
MY_EXTERN_C double abs_d(double x);
MY_EXTERN_C void quicksort(int* array, int n);
MY_EXTERN_C double get_random(double lowerbound, double upperbound);
MY_EXTERN_C void *realloc2 (void* ptr, size_t sizesbefore, size_t sizesnow);

/* graph construction and initialisation (core/paradmm.h:79-89) */
MY_EXTERN_C void startG(graph* G, int numdims);
MY_EXTERN_C void closeG(graph* G);
MY_EXTERN_C void printG(graph* G);

MY_EXTERN_C void initialize_RHOS_APHAS(graph* G, double rho0, double alpha0);
MY_EXTERN_C void initialize_X_N_Z_M_U_zeros(graph* G);
MY_EXTERN_C void initialize_X_N_Z_M_U_rand(graph* G, double X_l, double X_u, double N_l, double N_u,double Z_l, double Z_u, double M_l, double M_u, double U_l, double U_u);
MY_EXTERN_C void initialize_X_N_Z_M_U_optimal(graph* G, double Zinit[], double Uinit[], int maxiter);

MY_EXTERN_C void addNode(graph* G, void (*op)(double*, double*, double*, int, int, void*), void* operator_para, int opt_numVars, int* listvarix);
MY_EXTERN_C void addNodeWithParamSize(graph* G, void (*Operator)(double*, double*, double*, int, int, void*), void* operator_para, int operator_para_size, int opt_numvars, int* listvarix);

/* the update passes (core/paradmm.h:91-102) */
MY_EXTERN_C void update_outgoing_rhos(graph* G);

MY_EXTERN_C void updateXM(graph* G);
MY_EXTERN_C void updateX(graph* G);
MY_EXTERN_C void updateM(graph* G);

MY_EXTERN_C void updateZ(graph* G);

MY_EXTERN_C void updateUN(graph* G);
MY_EXTERN_C void updateU(graph* G);
MY_EXTERN_C void updateN(graph* G);

MY_EXTERN_C void updateG(graph* G);

/* The reference's OpenMP entry points (core/paradmm.h:108-115).  Here they
 * run on the GPU as well; numthreads is accepted and ignored.  They are
 * always exported; the prototypes stay behind OPENMP as in the reference. */
#ifdef OPENMP
MY_EXTERN_C void updateXOpenMP(graph* G);
MY_EXTERN_C void updateMOpenMP(graph* G);
MY_EXTERN_C void updateZOpenMP(graph* G);
MY_EXTERN_C void updateUOpenMP(graph* G);
MY_EXTERN_C void updateNOpenMP(graph* G);
MY_EXTERN_C void updateGOpenMP(graph* G, int maxiter, int numthreads);
#endif

#endif
