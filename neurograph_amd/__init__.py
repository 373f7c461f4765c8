"""neurograph_amd: the MI355X (gfx950) path for neurograph's all-pairs shortest paths and Brandes
betweenness centrality. The compute lives in libng_hip.so (HIP kernels behind a C ABI, see
include/neurograph_hip.h); this package is its Python binding."""

from ._lib import (  # noqa: F401
    NG_EDEVICE,
    NG_EINVAL,
    NG_ENOMEM,
    NG_EUNSUPPORTED,
    NG_LEN_ABS,
    NG_LEN_INVERSE,
    NG_LEN_RAW,
    NG_OK,
    Cohort,
    Context,
    Graph,
    HostGraph,
    contract_balls,
    NgError,
    generate_fixed_degree,
    generate_spatial,
    lib,
    load_csr,
    load_graph_dir,
    save_csr,
    version,
    write_graph_dir,
)
