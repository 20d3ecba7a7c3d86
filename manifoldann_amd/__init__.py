"""ManifoldANN.jl's data-parallel search core on MI355X (gfx950 HIP kernels).

The exact scan (BruteForceIndex), kNN-graph materialisation and the
MultiLevelIndex and IVFFlatIndex IVF paths (KMeansTransform, TopKRouting, list
scans, merge) run
in libmann_hip.so through its C ABI (include/mann.h); this package mirrors the
reference's Julia API on top of it.
"""
from ._lib import ArgumentError, DimensionMismatch, MannError, lib
from .index import (COSINE, EUCLIDEAN, SQEUCLIDEAN, BruteForceIndex, CosineDist, Dataset,
                    Euclidean, KNNGraph, Neighbor, SearchGraph, SqEuclidean, autotune,
                    build_index, build_knn_graph,
                    cosine_distance, default_distance, default_squared_distance,
                    euclidean_distance, graph_metadata, has_metadata, insert, knn_candidates,
                    last_emitted, last_stats,
                    neighbor_ids, node_metadata, query, query_arrays, recall_at_k,
                    symmetrize_adjacency)
from .ivf import HNSWIndex, IVFFlatIndex, MultiLevelIndex, build_ivf_hnsw_index

__all__ = [n for n in dir() if not n.startswith("_")]
