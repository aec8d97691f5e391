/*
 * Drop-in C++ entry points with the reference's exact signatures, built on
 * the C ABI in smg.h and exported (C++ linkage) by libsmg.so.
 *
 *   PESysMatGen   replaces PEGen_RayTracing_CircularHole/PESysMatGen.h:4-5
 *                 (defined PEGen_RayTracing_CircularHole/PESysMatGen.cu:1368-1437)
 *   scatter (9 args, combined crystal + collimator scatter)
 *                 replaces ScatterGen_RayTracing_CircularHole/scatter.h:4-5
 *                 (defined ScatterGen_RayTracing_CircularHole/scatter.cu:2013-2378)
 *   scatter (8 args, collimator scatter, r = 0.2 mm variant)
 *                 replaces ScatterGen_Collimator/scatter.h:4-5
 *                 (defined ScatterGen_Collimator/scatter.cu:1511-1747)
 *   scatter (7 args, crystal scatter)
 *                 replaces ScatterGen_Crystal/scatter.h:4
 *                 (defined ScatterGen_Crystal/scatter.cu:1528-1738)
 *
 * Same argument meaning: parameter arrays of 80000/80000/100/100 floats,
 * img[20] = rotation index (a non-negative whole number: any other value
 * prints a message and exits), dst receives the N x V block of that rotation,
 * phys[8]/phys[9] == 1 computes a geometry bitmap and writes it to the named
 * cache file, otherwise the file is read.  Returns the number of voxels.
 *
 * Behaviour kept from the reference: a device index beyond the device count
 * runs on device 0 with a message (ScatterGen_RayTracing_CircularHole/
 * scatter.cu:2034-2037).  Changed: a failure (unreadable cache file, HIP
 * error, geometry beyond a capacity) prints a message and exits with
 * EXIT_FAILURE instead of exit(0) / exit(hip error).  The combined scatter()
 * sums exactly N x V elements (the reference's sum loop runs over N x V x R,
 * scatter.cu:2354-2357).
 *
 * Environment: SMG_PE_SUBSAMPLES="sx,sy,sz" overrides the PE sub-cells
 * (default 8,16,8, PESysMatGen.cu:1037); SMG_SCATTER_SUBSAMPLES likewise
 * for the scatter detection cells (default 1,1,1); SMG_PRECISION=fp64 selects
 * the all-FP64 kernels.
 */
#ifndef SMG_REFERENCE_H
#define SMG_REFERENCE_H

int PESysMatGen(float* parameter_Collimator, float* parameter_Detector, float* parameter_Image,
                float* dst, int cuda_id);

int scatter(float* parameter_Collimator, float* parameter_Detector, float* parameter_Image,
            float* parameter_Physics, float* PE_SysMat, const char* FnameGeoCrystal,
            const char* FnameGeoCollimator, float* dst, int cuda_id);

int scatter(float* parameter_Collimator, float* parameter_Detector, float* parameter_Image,
            float* parameter_Physics, float* PE_SysMat, const char* FnameGeo, float* dst,
            int cuda_id);

int scatter(float* parameter_Detector, float* parameter_Image, float* parameter_Physics,
            float* PE_SysMat, const char* FnameGeo, float* dst, int cuda_id);

/* Extension with no reference counterpart, for the drivers: the scatter block
 * of which = 1 (crystal, as the 7-argument scatter), 2 (collimator r = 0.2 mm,
 * as the 8-argument one) or 3 (combined, as the 9-argument one) into
 * scatter_dst, and PE_add + scatter into sum_dst, both formed on the GPU
 * (smg_scatter_sum); the reference drivers add PE and scatter on the host
 * (ScatterGen_RayTracing_CircularHole/ScatterGen_CircularHole.cpp:211-230).
 * Either output may be NULL, not both; parameter_Collimator and
 * FnameGeoCollimator are unused for which = 1, FnameGeoCrystal for which = 2. */
int scatter_and_sum(int which, float* parameter_Collimator, float* parameter_Detector,
                    float* parameter_Image, float* parameter_Physics, float* PE_SysMat, float* PE_add,
                    const char* FnameGeoCrystal, const char* FnameGeoCollimator, float* scatter_dst,
                    float* sum_dst, int cuda_id);

#endif /* SMG_REFERENCE_H */
