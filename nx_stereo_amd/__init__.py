"""nx_stereo_amd: NxStereo's semi-global matching path and the stages around it on MI355X.

The product is libnxs_hip.so (include/nxs_hip.h), hand-written HIP for gfx950
behind the C ABI a NIF binds; this package holds its Python bindings
(`_native`, `segment.HipSegment`), the protocol records, the stages'
settings API, the build script and the seeded synthetic inputs.
"""
