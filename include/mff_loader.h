/* Mufflon's scene-file loader (the mffloader plugin) for the HIP path tracer:
 * reads a scene in the Mufflon file format (a JSON description plus a binary
 * .mff file with the geometry and instances) and builds the world through the
 * core C API of libmufflon_hip.so (include/mufflon_core.h).
 *
 * Library: mufflon_amd/lib/libmffloader.so. The entry points keep the names,
 * arguments and return conventions of the reference's loader interface
 * (src/mffloader/interface/mff_interface.h:15-44), so front ends that load
 * mffloader.dll (bindings.py:313-317) bind them unchanged.
 *
 * What the HIP path tracer cannot render is reported, not approximated: a
 * scenario that uses a focus or orthographic camera, a spot, directional,
 * sky or goniometric light, or a Walter, Oren-Nayar, blend, fresnel, alpha-
 * or displacement-mapped material is not created (with a warning); if that is
 * the default scenario, or a LoD holds quads or spheres, loading fails with a
 * message in loader_get_dll_error(). Custom vertex/face attributes and bone
 * animation data are read past. */
#ifndef MFF_LOADER_H
#define MFF_LOADER_H

#include <stddef.h>
#include <stdint.h>

#include "mufflon_core.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef void* MufflonLoaderInstanceHdl;

/* mff_interface.h:15-19 */
typedef enum { LOADER_SUCCESS, LOADER_ERROR, LOADER_ABORT } LoaderStatus;

/* Message of the last failed call (mff_interface.h:21). The string stays
 * valid until the next loader call fails on this thread. */
const char* loader_get_dll_error(void);
/* Which loader messages reach stderr (LOG_WARNING by default; interface.cpp:50-72). */
Boolean loader_set_log_level(LogLevel level);
/* A loader bound to one core instance (mff_interface.h:22-23). */
MufflonLoaderInstanceHdl loader_initialize(MufflonInstanceHdl mffInstHdl);
void loader_destroy(MufflonLoaderInstanceHdl hdl);
/* Clears the world, reads the JSON file and its binary file, creates
 * objects, instances, cameras, lights, materials and scenarios, registers the
 * on-demand LoD loader with the core and loads the default scenario
 * (interface.cpp:97-129, json_loader.cpp:1031-1185). LOADER_ABORT after
 * loader_abort() from another thread. */
LoaderStatus loader_load_json(MufflonLoaderInstanceHdl hdl, const char* path);
/* Not provided by this loader: always LOADER_ERROR (mff_interface.h:25). */
LoaderStatus loader_save_scene(MufflonLoaderInstanceHdl hdl, const char* path);
/* Loads LoD `lod` (clamped to the object's last one) of an object of the
 * loaded file (mff_interface.h:26, binary.cpp:672-744, 901-978). The core
 * calls it through mufflon_set_lod_loader. */
Boolean loader_load_lod(MufflonLoaderInstanceHdl hdl, ObjectHdl obj, uint32_t lod, Boolean asReduced);
/* The sorted unique material indices of an object (file version >= 1.7; 0
 * read otherwise) (mff_interface.h:27-28, binary.cpp:980-1045). */
Boolean loader_load_object_material_indices(MufflonLoaderInstanceHdl hdl, const uint32_t objId,
                                            uint16_t* indexBuffer, uint32_t* readIndices);
/* Sizes of every LoD of every object, objects in file order (mff_interface.h:29). */
Boolean loader_load_lod_metadatas(MufflonLoaderInstanceHdl hdl, LodMetadata* data, size_t* read);
/* Asks a running loader_load_json to stop (mff_interface.h:30). */
Boolean loader_abort(MufflonLoaderInstanceHdl hdl);
/* The current loading stage as text (mff_interface.h:31). */
const char* loader_get_loading_status(MufflonLoaderInstanceHdl hdl);
void loader_profiling_enable(void);
void loader_profiling_disable(void);

#ifdef __cplusplus
}
#endif

#endif /* MFF_LOADER_H */
