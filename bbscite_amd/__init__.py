"""bbscite_amd: MI355X (gfx950) implementation of bbSCITE's MCMC
mutation-tree search behind a C-ABI (include/bbscite.h).

The compute path is the in-tree HIP library libbbscite.so; this package
only binds it (ctypes) and builds it. There is no CPU fallback.
"""
from ._capi import (  # noqa: F401
    Context,
    Params,
    Server,
    server_config,
    BbsError,
    default_params,
    encode_matrix,
    generate_dataset,
    initial_tree,
    lib,
    newick_to_parent_vector,
    score_attachments,
    tost,
    parent_vector_to_matrix,
    parse_csv,
    pruefer_to_parent_vector,
    random_trees,
    tree_to_graphviz,
    tree_to_newick,
)
