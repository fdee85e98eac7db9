/*
 * yasps_amd - C ABI of a run-time generated user energy module.
 *
 * yasps_amd/usercode.py differentiates a user's SymPy energy, emits its element function
 * and builds a module (a shared library linked against libyasps_amd.so) that runs the
 * library's element kernels (csrc/energy_kernels.hpp) for that one energy. Each module
 * exports these entry points; they replace the reference's generated element kernels
 * (codeGenerator/codeGenerator.pyx:159-369 emits the element functions that
 * kernel/Hessian/hessianKernelNoProject.pyx:16-264 and the projecting variants call).
 */
#ifndef YASPS_AMD_USER_ENERGY_H
#define YASPS_AMD_USER_ENERGY_H

#include "yasps_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Element shape of the module's energy: vertices per element, values per vertex, element
 * parameters per element, scene parameter slots (at least one: an energy without scene
 * parameters reads one unused slot). As ysp_energy_arity for a library energy. */
int ysp_user_energy_arity(uint32_t* arity, uint32_t* width, uint32_t* element_params,
                          uint32_t* scene_params);

/* Energy gradient and (projected, method 0 / 1 / 2) Hessian of every element, scattered
 * as ysp_eval_element_energy does for a library energy (same arguments without the
 * kind). */
int ysp_user_energy_eval(const ysp_union_f64* positions, const ysp_affine_children* affine,
                         const uint32_t* elem2v, uint32_t n, const double* element_params,
                         const double* scene_params, uint32_t projection_method,
                         const ysp_scatter_args* out, void* stream);

/* Element blocks of d g / d theta from the caller's d p / d theta and d s / d theta, as
 * ysp_element_param_jacobian for a library energy (same arguments without the kind):
 * every element parameter and scene slot of the energy has its generated derivative. */
int ysp_user_energy_param_jacobian(const ysp_union_f64* positions,
                                   const uint32_t* element_vertices, uint32_t n,
                                   const double* element_params, const double* scene_params,
                                   uint32_t scene_stride, const double* dparams,
                                   uint32_t dparams_stride, const double* dscene,
                                   uint32_t dscene_stride, uint32_t w, double* blocks,
                                   void* stream);

/* energy[e] of every element (line search), as ysp_element_energy_values. */
int ysp_user_energy_values(const ysp_union_f64* positions, const uint32_t* elem2v, uint32_t n,
                           const double* element_params, const double* scene_params,
                           double* energy, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* YASPS_AMD_USER_ENERGY_H */
