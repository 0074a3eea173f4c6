/* Exports of the VapourSynth plugin shells over libsvpf:
 *   libsvpflow1_vs.so  (svp1: Super, Analyse)   replaces the reference's svpflow1 plugin
 *   libsvpflow2_vs.so  (svp2: SmoothFps)        replaces the reference's svpflow2 plugin
 * Both export exactly these two symbols (svpflow1_vs.def:1-4, svpflow2_vs.def:1-4).
 * Types are VapourSynth API 3's, passed as opaque pointers here; the plugins read the
 * VSAPI table by the byte offsets the reference binds (open-svpflow_amd/vsplugin/vsapi3.h). */
#ifndef SVPFLOW_VS_H
#define SVPFLOW_VS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* VapourSynthPluginInit (crates/svpflow1/src/abi.rs:10-46, crates/svpflow2/src/abi.rs):
 * config(identifier, namespace, name, 0x30002, readonly 1, plugin), then one
 * register(name, args, create, NULL, plugin) per filter. */
typedef void (*svpvs_config_plugin)(const char* identifier, const char* ns, const char* name, int api_version,
                                    int readonly, void* plugin);
typedef void (*svpvs_register_function)(const char* name, const char* args, void* create, void* user,
                                        void* plugin);
void VapourSynthPluginInit(svpvs_config_plugin config, svpvs_register_function reg, void* plugin);

/* svpGetVersion (crates/svpflow1/src/abi.rs:3-8): 1174405393. */
int64_t svpGetVersion(void);

#ifdef __cplusplus
}
#endif
#endif /* SVPFLOW_VS_H */
