/* sysmat_ref.h -- the reference's C++-linkage entry points, exported by
 * libsysmat_hip.so with the reference's exact signatures so that the
 * reference's own driver programs link against this library unchanged.
 * Each forwards to the C ABI in sysmat.h and returns its value (a negative
 * SM_E* code on failure, after printing sm_last_error() to stderr).
 *
 *   PESysMatGen   PEGen_RayTracing_CircularHole/PESysMatGen.h:4-9      -> sm_PESysMatGen
 *   scatter (9)   ScatterGen_RayTracing_CircularHole/scatter.h:4-5     -> sm_scatter
 *   scatter (8)   ScatterGen_Collimator/scatter.h:4-5                  -> sm_scatter_collimator_only
 *   scatter (7)   ScatterGen_Crystal/scatter.h:4                       -> sm_scatter_crystal_only
 */
#ifndef SYSMAT_REF_H_
#define SYSMAT_REF_H_
#ifdef __cplusplus

int PESysMatGen(float* parameter_Collimator, float* parameter_Detector, float* parameter_Image, float* dst,
                int cuda_id);

int scatter(float* parameter_Collimator, float* parameter_Detector, float* parameter_Image, float* parameter_Physics,
            float* PE_SysMat, const char* FnameGeoCrystal, const char* FnameGeoCollimator, float* dst, int cuda_id);

int scatter(float* parameter_Collimator, float* parameter_Detector, float* parameter_Image, float* parameter_Physics,
            float* PE_SysMat, const char* FnameGeo, float* dst, int cuda_id);

int scatter(float* parameter_Detector, float* parameter_Image, float* parameter_Physics, float* PE_SysMat,
            const char* FnameGeo, float* dst, int cuda_id);

#endif /* __cplusplus */
#endif /* SYSMAT_REF_H_ */
