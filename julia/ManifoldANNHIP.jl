"""
    ManifoldANNHIP

Binds ManifoldANN.jl's data-parallel search path to the MI355X library
`libmann_hip.so` (C ABI in `include/mann.h`) by adding more specific methods of
the reference's own functions, so callers (and the Python harness's
`jl.query(index, data, Q, k)`, benchmarking/wrappers/manifoldann.py:231-237)
keep calling them unchanged:

- `query(::BruteForceIndex{Float32,D}, data, Q | q, k)`  (bruteforce.jl:46-71,
  ann_index.jl:71-101)
- `build_knn_graph(::BruteForceIndex{Float32,D}, data; k, include_self,
  directed, metadata)`  (knn_graph.jl:42-81), and the candidate re-query of
  `build_weighted_graph` (weighted_knn_graph.jl:489-515) as one call
- `fit!(::KMeansTransform{Euclidean,Float32}, X; rng)`  (KMeansTransform.jl:106-162)
- `query(::MultiLevelIndex{<:KMeansTransform,<:BruteForceIndex}, data, Q | q, k)`
  (multilevel/query.jl:25-39, :279-300) with the index's merge strategy
- `query(::IVFFlatIndex{Float32}, data, Q | q, k; nprobe)`  (ivf_flat.jl:138-168)
- `query(::MultiLevelIndex{<:KMeansTransform,<:HNSWIndex}, data, Q | q, k; ef_search)`
  (build_ivf_hnsw_index, multilevel/ivf_hnsw.jl:46-88): the children's own
  graphs (their HNSWLayer slabs) are installed on the device on first query and
  searched there

`D` is Euclidean, SqEuclidean or CosineDist: the metric comes from the index.
Other element types, metrics and index configurations keep the reference
methods (these methods are simply not more specific for them).

Data ownership follows the reference (ann_index.jl:1-10): indices never store
`data`.  A call borrows `data` for its duration and prepares a device copy for
that call only, unless the caller has attached the matrix with `attach!`, which
keeps one device copy per (matrix, metric family) until `detach!` or until the
matrix is garbage collected.  An attached matrix must not be mutated in place
(detach and attach again).  The IVF indexes keep their list-contiguous device
layout in a registry keyed by the index's own mutable parts (its transform or
its lists), built on first query; `release!(index)` frees it.

Randomness: `fit!` draws one UInt64 seed from `rng` and the library's kmeans++
and subsample draws follow splitmix64 from it (Julia's RNG streams are not
reproduced on the GPU).
"""
module ManifoldANNHIP

using ManifoldANN
using ManifoldANN: Neighbor, BruteForceIndex, KMeansTransform, MultiLevelIndex, IVFFlatIndex,
                   KNNGraph, SimpleMerge, DisjointMerge, TopKRouting, HNSWIndex, max_degree
using Distances: Euclidean, SqEuclidean, CosineDist
using Random: AbstractRNG, default_rng
import ManifoldANN: query, build_knn_graph, fit!

export attach!, detach!, release!

const LIB = get(ENV, "MANN_HIP_LIB",
                joinpath(@__DIR__, "..", "manifoldann_amd", "libmann_hip.so"))

const HIPMetric = Union{Euclidean,SqEuclidean,CosineDist}
metric_code(::Euclidean) = Cint(0)
metric_code(::SqEuclidean) = Cint(1)
metric_code(::CosineDist) = Cint(2)

# Status codes -> the exception the reference throws at the same point.
function check(rc::Cint)
    rc == 0 && return nothing
    msg = unsafe_string(ccall((:mann_last_error, LIB), Cstring, ()))
    rc == -1 && throw(DimensionMismatch(msg))
    rc == -2 && throw(ArgumentError(msg))
    rc == -3 && throw(AssertionError(msg))
    error("libmann_hip status $rc: $msg")
end

# Columns must be contiguous (a d x n matrix or a column range of one).
dense_cols(X::StridedMatrix{Float32}) = stride(X, 1) == 1 ? X : collect(X)
dense_vec(x::StridedVector{Float32}) = stride(x, 1) == 1 ? x : collect(x)

# ---------------------------------------------------------------------------
# device copies of `data`
# ---------------------------------------------------------------------------

mutable struct DeviceDataset
    h::Ptr{Cvoid}
    cosine::Bool
    ptr::Ptr{Float32}
    dims::Tuple{Int,Int}
end

function DeviceDataset(X::StridedMatrix{Float32}, cosine::Bool)
    stride(X, 1) == 1 || throw(ArgumentError("data columns must be contiguous"))
    h = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve X check(ccall((:mann_dataset_create, LIB), Cint,
                               (Ptr{Float32}, Int64, Int64, Int64, Cint, Ref{Ptr{Cvoid}}),
                               X, size(X, 1), size(X, 2), stride(X, 2),
                               cosine ? Cint(2) : Cint(0), h))
    dd = DeviceDataset(h[], cosine, pointer(X), size(X))
    finalizer(release!, dd)
end

function release!(dd::DeviceDataset)
    if dd.h != C_NULL
        ccall((:mann_dataset_destroy, LIB), Cint, (Ptr{Cvoid},), dd.h)
        dd.h = C_NULL
    end
    return nothing
end

# matrix -> its attached device copies (at most one per metric family)
const ATTACHED = WeakKeyDict{Matrix{Float32},Vector{DeviceDataset}}()
const LOCK = ReentrantLock()

"""
    attach!(data::Matrix{Float32}; distance=Euclidean()) -> data

Keep a device copy of `data` for the metric family of `distance` (CosineDist
uses unit-normalised bf16 rows, the L2 metrics share one copy) until
`detach!(data)`.  Calls that pass this matrix use the copy instead of
preparing one per call.
"""
function attach!(data::Matrix{Float32}; distance::HIPMetric = Euclidean())
    cosine = distance isa CosineDist
    lock(LOCK) do
        list = get!(ATTACHED, data, DeviceDataset[])
        any(dd -> dd.cosine == cosine, list) || push!(list, DeviceDataset(data, cosine))
    end
    return data
end

function detach!(data::Matrix{Float32})
    lock(LOCK) do
        for dd in get(ATTACHED, data, DeviceDataset[])
            release!(dd)
        end
        delete!(ATTACHED, data)
    end
    return data
end

# Run f(dd) with a device copy of `data` for `metric`: the attached one, or
# one prepared for this call and freed after it.
function with_dataset(f, data::StridedMatrix{Float32}, metric::HIPMetric)
    cosine = metric isa CosineDist
    if data isa Matrix{Float32}
        dd = lock(LOCK) do
            for d in get(ATTACHED, data, DeviceDataset[])
                d.cosine == cosine && d.ptr == pointer(data) && d.dims == size(data) && return d
            end
            nothing
        end
        dd === nothing || return GC.@preserve data f(dd)
    end
    X = dense_cols(data)
    dd = DeviceDataset(X, cosine)
    try
        return GC.@preserve X f(dd)
    finally
        release!(dd)
    end
end

function to_neighbors(ids::Matrix{Int64}, dist::Matrix{Float32})
    k, nq = size(ids)
    return [Neighbor{Float32}[Neighbor{Float32}(ids[r, q], dist[r, q]) for r in 1:k
                              if ids[r, q] != 0] for q in 1:nq]
end

# ---------------------------------------------------------------------------
# BruteForceIndex
# ---------------------------------------------------------------------------

function bf_knn(index::BruteForceIndex{Float32,D}, data::StridedMatrix{Float32},
                Q::StridedMatrix{Float32}, k::Int) where {D<:HIPMetric}
    nq = size(Q, 2)
    ids = zeros(Int64, k, nq)
    dist = fill(Inf32, k, nq)
    Qc = dense_cols(Q)
    with_dataset(data, index.distance) do dd
        GC.@preserve Qc ids dist check(ccall((:mann_bf_knn, LIB), Cint,
            (Ptr{Cvoid}, Ptr{Float32}, Int64, Int64, Int64, Int64, Cint, Ptr{Int64},
             Ptr{Float32}),
            dd.h, Qc, size(Qc, 1), nq, stride(Qc, 2), k, metric_code(index.distance), ids,
            dist))
    end
    return to_neighbors(ids, dist)
end

# batch (ann_index.jl:71-101): the dimension check holds for empty batches too
function query(index::BruteForceIndex{Float32,D}, data::StridedMatrix{Float32},
               queries::StridedMatrix{Float32}, k::Integer) where {D<:HIPMetric}
    size(queries, 1) == size(data, 1) || throw(DimensionMismatch(
        "Expected queries with $(size(data, 1)) rows, got $(size(queries, 1))"))
    size(queries, 2) == 0 && return Vector{Vector{Neighbor{Float32}}}()
    ManifoldANN.validate_index_query_matrix(index, queries)
    size(data, 2) >= index.n_points || throw(ArgumentError(
        "Data contains $(size(data, 2)) points but index tracks $(index.n_points)"))
    k <= 0 && return [Neighbor{Float32}[] for _ in 1:size(queries, 2)]
    return bf_knn(index, data, queries, Int(k))
end

# one query (bruteforce.jl:46-71)
function query(index::BruteForceIndex{Float32,D}, data::StridedMatrix{Float32},
               q::StridedVector{Float32}, k::Integer) where {D<:HIPMetric}
    ManifoldANN.validate_index_dimensions(index, data, q)
    kk = min(Int(k), size(data, 2))
    kk <= 0 && return Neighbor{Float32}[]
    return bf_knn(index, data, reshape(dense_vec(q), :, 1), kk)[1]
end

# build_knn_graph (knn_graph.jl:42-81): the self-join on the GPU, self removed
# by id, the error of _prepare_neighbors when a row comes back short, and the
# undirected union (_symmetrize_adjacency, :196-210) on the GPU
function build_knn_graph(index::BruteForceIndex{Float32,D}, data::StridedMatrix{Float32};
                         k::Union{Nothing,Integer} = nothing, include_self::Bool = false,
                         directed::Bool = true,
                         metadata::Union{Nothing,AbstractVector} = nothing) where {D<:HIPMetric}
    n = size(data, 2)
    n > 0 || throw(ArgumentError("Dataset must contain at least one point"))
    kk = ManifoldANN._resolve_graph_k(index, k)
    maxnb = n - (include_self ? 0 : 1)
    (0 < kk <= maxnb) || throw(ArgumentError(
        "k must satisfy 0 < k <= $maxnb for dataset with $n points"))
    ids = zeros(Int64, kk, n)
    with_dataset(data, index.distance) do dd
        GC.@preserve ids check(ccall((:mann_knn_graph, LIB), Cint,
            (Ptr{Cvoid}, Int64, Cint, Cint, Ptr{Int64}, Ptr{Float32}),
            dd.h, kk, Cint(include_self), metric_code(index.distance), ids, C_NULL))
    end
    for i in 1:n
        c = count(!iszero, @view ids[:, i])
        c == kk || throw(ArgumentError(
            "Index could only supply $c usable neighbors for point $i. " *
            "Consider lowering k or setting include_self=true."))
    end
    original_k = kk
    if directed
        adjacency = [Vector{Int}(ids[:, i]) for i in 1:n]
    else
        off = zeros(Int64, n + 1)
        flat = zeros(Int64, 2 * n * kk)
        total = Ref{Int64}(0)
        maxdeg = Ref{Int64}(0)
        GC.@preserve ids off flat check(ccall((:mann_knn_graph_symmetrize, LIB), Cint,
            (Ptr{Int64}, Int64, Int64, Ptr{Int64}, Ptr{Int64}, Ref{Int64}, Ref{Int64}),
            ids, n, kk, off, flat, total, maxdeg))
        adjacency = [Vector{Int}(flat[off[i]+1:off[i+1]]) for i in 1:n]
        kk = Int(maxdeg[])
    end
    meta = ManifoldANN._create_graph_metadata(original_k, directed, metadata, n)
    return KNNGraph(adjacency, kk, include_self, meta)
end

"""
    knn_candidates(index, data, candidate_k) -> Vector{Vector{Int}}

The candidate lists of `build_weighted_graph` with `candidate_k > k`
(weighted_knn_graph.jl:497-503) in one GPU call: every point queried with
`candidate_k + 1`, its own id dropped, the first `candidate_k` kept; short
lists are kept as they are.
"""
function knn_candidates(index::BruteForceIndex{Float32,D}, data::StridedMatrix{Float32},
                        candidate_k::Integer) where {D<:HIPMetric}
    n = size(data, 2)
    ids = zeros(Int64, candidate_k, n)
    cnt = zeros(Int64, n)
    with_dataset(data, index.distance) do dd
        GC.@preserve ids cnt check(ccall((:mann_knn_candidates, LIB), Cint,
            (Ptr{Cvoid}, Int64, Cint, Ptr{Int64}, Ptr{Float32}, Ptr{Int64}),
            dd.h, candidate_k, metric_code(index.distance), ids, C_NULL, cnt))
    end
    return [Vector{Int}(ids[1:cnt[i], i]) for i in 1:n]
end

# the per-point candidate queries of the unshared geometry fit
# (weighted_knn_graph.jl:489-515) as one GPU call; the fits are unchanged
function ManifoldANN._fit_geometries_with_candidates(::ManifoldANN.NoSharing,
        method::ManifoldANN.AbstractLocalGeometryMethod, graph::KNNGraph,
        index::BruteForceIndex{Float32,D}, data::StridedMatrix{Float32},
        candidate_k::Integer) where {D<:HIPMetric}
    n = length(graph)
    n == 0 && return ManifoldANN.AbstractLocalGeometry[]
    cands = knn_candidates(index, data, candidate_k)
    g1 = ManifoldANN.fit_geometry(method, data, 1, cands[1]; graph = graph)
    geometries = Vector{typeof(g1)}(undef, n)
    geometries[1] = g1
    Threads.@threads for i in 2:n
        geometries[i] = ManifoldANN.fit_geometry(method, data, i, cands[i]; graph = graph)
    end
    return geometries
end

# ---------------------------------------------------------------------------
# KMeansTransform
# ---------------------------------------------------------------------------

# fit! (KMeansTransform.jl:106-162): the optional subsample (max(subsample_size,
# k) distinct points in the order drawn), the initialisation, Lloyd and, after a
# subsampled fit, the full-data E-step all run in the library
function fit!(t::KMeansTransform{Euclidean,Float32}, X::Matrix{Float32};
              rng::AbstractRNG = default_rng())
    d, n = size(X)
    @assert t.k <= n "Cannot fit $(t.k) clusters with only $n points"
    init = t.init == :kmeans_plus_plus ? Cint(1) :
           t.init == :random ? Cint(2) :
           throw(ArgumentError("Unknown initialization strategy: $(t.init)"))
    seed = rand(rng, UInt64)
    sub = (t.subsample_size !== nothing && t.subsample_size < n) ? Int64(t.subsample_size) : 0
    C = zeros(Float32, d, t.k)
    assign = zeros(Int64, n)
    iters = Ref{Int64}(0)
    with_dataset(X, Euclidean()) do dd
        GC.@preserve C assign check(ccall((:mann_kmeans_fit, LIB), Cint,
            (Ptr{Cvoid}, Int64, Int64, Float64, Cint, UInt64, Ptr{Float32}, Int64,
             Ptr{Float32}, Ptr{Int64}, Ref{Int64}),
            dd.h, t.k, t.max_iters, t.tol, init, seed, C_NULL, sub, C, assign, iters))
    end
    t.centroids = C
    t.centroid_norms = vec(sum(abs2, C; dims = 1))
    t.pending_assignments = Vector{Int}(assign)
    return t
end

# ---------------------------------------------------------------------------
# MultiLevelIndex{KMeansTransform, BruteForceIndex} and IVFFlatIndex
# ---------------------------------------------------------------------------

mutable struct IVFHandle
    h::Ptr{Cvoid}
    data_ptr::Ptr{Float32}
    dims::Tuple{Int,Int}
    cosine::Bool
    graphs::Bool    # HNSW terminals installed
end

function release!(h::IVFHandle)
    if h.h != C_NULL
        ccall((:mann_ivf_destroy, LIB), Cint, (Ptr{Cvoid},), h.h)
        h.h = C_NULL
    end
    return nothing
end

# keyed by a mutable part of each index (its KMeansTransform, or an IVFFlatIndex's
# list vector), so the device layout lives as long as the index
const IVF = WeakKeyDict{Any,IVFHandle}()

const HIPMultiLevel = MultiLevelIndex{<:KMeansTransform{Euclidean},<:BruteForceIndex}

ivf_key(index::MultiLevelIndex) = index.root.transform
ivf_key(index::IVFFlatIndex) = index.lists

# point -> 1-based bucket from the index's own partition
function bucket_assignment(index::MultiLevelIndex, n::Int)
    root = index.root
    a = zeros(Int64, n)
    lookup = root.bucket_lookup
    for (b, child) in enumerate(lookup === nothing ? eachindex(root.indices) : lookup)
        child == 0 && continue
        for id in root.id_mappings[child]
            a[id] = b
        end
    end
    return a
end

function bucket_assignment(index::IVFFlatIndex, n::Int)
    a = zeros(Int64, n)
    for (c, ids) in enumerate(index.lists), id in ids
        a[id] = c
    end
    return a
end

function ivf_handle(index, data::StridedMatrix{Float32}, metric::HIPMetric)
    key = ivf_key(index)
    cosine = metric isa CosineDist
    lock(LOCK) do
        h = get(IVF, key, nothing)
        if h !== nothing && h.h != C_NULL && h.data_ptr == pointer(data) &&
           h.dims == size(data) && h.cosine == cosine
            return h
        end
        h === nothing || release!(h)
        n = size(data, 2)
        assign = bucket_assignment(index, n)
        C = index isa IVFFlatIndex ? index.centroids : index.root.transform.centroids
        out = Ref{Ptr{Cvoid}}(C_NULL)
        with_dataset(data, metric) do dd
            GC.@preserve C assign if index isa IVFFlatIndex
                check(ccall((:mann_ivf_flat_build, LIB), Cint,
                            (Ptr{Cvoid}, Ptr{Float32}, Int64, Ptr{Int64}, Cint, Ref{Ptr{Cvoid}}),
                            dd.h, C, size(C, 2), assign, metric_code(index.centroid_metric), out))
            else
                check(ccall((:mann_ivf_build, LIB), Cint,
                            (Ptr{Cvoid}, Ptr{Float32}, Int64, Ptr{Int64}, Ref{Ptr{Cvoid}}),
                            dd.h, C, size(C, 2), assign, out))
            end
        end
        h = IVFHandle(out[], pointer(data), size(data), cosine, false)
        finalizer(release!, h)
        IVF[key] = h
        return h
    end
end

release!(index::Union{MultiLevelIndex,IVFFlatIndex}) =
    lock(LOCK) do
        h = pop!(IVF, ivf_key(index), nothing)
        h === nothing || release!(h)
        nothing
    end

function ivf_search(index, data::StridedMatrix{Float32}, Q::StridedMatrix{Float32}, k::Int,
                    nprobe::Int, metric::HIPMetric, merge::Cint, probe_order::Bool)
    nq = size(Q, 2)
    ids = zeros(Int64, k, nq)
    dist = fill(Inf32, k, nq)
    Qc = dense_cols(Q)
    h = ivf_handle(index, data, metric)
    GC.@preserve Qc ids dist check(ccall((:mann_ivf_search_ex, LIB), Cint,
        (Ptr{Cvoid}, Ptr{Float32}, Int64, Int64, Int64, Int64, Int64, Cint, Ptr{Float32}, Cint,
         Cint, Ptr{Int64}, Ptr{Float32}),
        h.h, Qc, size(Qc, 1), nq, stride(Qc, 2), k, nprobe, metric_code(metric), C_NULL, merge,
        Cint(probe_order), ids, dist))
    return to_neighbors(ids, dist)
end

merge_code(m) = m isa DisjointMerge ? Cint(1) : Cint(0)
routing_k(index::MultiLevelIndex) = index.root.routing_strategy::TopKRouting

# batched MultiLevelIndex query (query.jl:279-300): per-list results merged in
# child (bucket) order
function query(index::HIPMultiLevel, data::StridedMatrix{Float32}, Q::StridedMatrix{Float32},
               k::Integer)
    size(Q, 1) == size(data, 1) || throw(DimensionMismatch(
        "Expected queries with $(size(data, 1)) rows, got $(size(Q, 1))"))
    size(Q, 2) == 0 && return Vector{Vector{Neighbor{Float32}}}()
    k <= 0 && return [Neighbor{Float32}[] for _ in 1:size(Q, 2)]
    return ivf_search(index, data, Q, Int(k), routing_k(index).k, index.distance,
                      merge_code(index.merge_strategy), false)
end

# one query (query.jl:25-106): per-list results merged in probe order
function query(index::HIPMultiLevel, data::StridedMatrix{Float32}, q::StridedVector{Float32},
               k::Integer)
    length(q) == size(data, 1) || throw(DimensionMismatch(
        "Expected query with $(size(data, 1)) rows, got $(length(q))"))
    k <= 0 && return Neighbor{Float32}[]
    return ivf_search(index, data, reshape(dense_vec(q), :, 1), Int(k), routing_k(index).k,
                      index.distance, merge_code(index.merge_strategy), true)[1]
end

# IVFFlatIndex (ivf_flat.jl:95-168): nprobe clamped to [1, nlist], bounded-heap merge
function query(index::IVFFlatIndex{Float32,Dm,Dc}, data::StridedMatrix{Float32},
               Q::StridedMatrix{Float32}, k::Integer;
               nprobe::Int = index.default_nprobe) where {Dm<:HIPMetric,Dc<:HIPMetric}
    size(data, 1) == index.dimension ||
        throw(DimensionMismatch("Expected data with $(index.dimension) rows"))
    size(Q, 2) == 0 && return Vector{Vector{Neighbor{Float32}}}()
    size(Q, 1) == index.dimension ||
        throw(DimensionMismatch("Expected query of length $(index.dimension)"))
    k <= 0 && return [Neighbor{Float32}[] for _ in 1:size(Q, 2)]
    np = clamp(nprobe, 1, length(index.lists))
    return ivf_search(index, data, Q, Int(k), np, index.distance, Cint(0), true)
end

function query(index::IVFFlatIndex{Float32,Dm,Dc}, data::StridedMatrix{Float32},
               q::StridedVector{Float32}, k::Integer;
               nprobe::Int = index.default_nprobe) where {Dm<:HIPMetric,Dc<:HIPMetric}
    ManifoldANN._validate_dimensions(index, data, length(q))
    k <= 0 && return Neighbor{Float32}[]
    np = clamp(nprobe, 1, length(index.lists))
    return ivf_search(index, data, reshape(dense_vec(q), :, 1), Int(k), np, index.distance,
                      Cint(0), true)[1]
end

# ---------------------------------------------------------------------------
# MultiLevelIndex{KMeansTransform, HNSWIndex} (build_ivf_hnsw_index)
# ---------------------------------------------------------------------------

const HIPMultiLevelHNSW = MultiLevelIndex{<:KMeansTransform{Euclidean},<:HNSWIndex}

# The children's HNSWLayer slabs (hnsw/types.jl:18-49) as one [layer][row][slot]
# array over the handle's list-contiguous rows (buckets ascending, each child's
# local ids in order), 0-based local ids, installed with mann_ivf_hnsw_set_graph
# (which checks every id and degree before installing).
function install_graphs!(h::IVFHandle, index::MultiLevelIndex, n::Int)
    root = index.root
    kids = root.indices
    L = maximum(c.max_layer for c in kids) + 1
    M = maximum(max_degree(c.neighbor_policy) for c in kids)
    nbr = zeros(Int32, M + 1, n, L)
    deg = zeros(Int32, n, L)
    nb = length(root.bucket_lookup)
    entry = fill(Int32(-1), nb)
    maxl = fill(Int32(-1), nb)
    row = 0
    for b in 1:nb
        c = root.bucket_lookup[b]
        c == 0 && continue
        child = kids[c]
        m = length(root.id_mappings[c])
        for l in 0:child.max_layer
            layer = child.layers[l + 1]
            for i in 1:m
                dg = layer.degree[i]
                deg[row + i, l + 1] = dg
                for j in 1:dg
                    nbr[j, row + i, l + 1] = layer.neighbors[j, i] - 1
                end
            end
        end
        entry[b] = child.entry_point - 1
        maxl[b] = child.max_layer
        row += m
    end
    GC.@preserve nbr deg entry maxl check(ccall((:mann_ivf_hnsw_set_graph, LIB), Cint,
        (Ptr{Cvoid}, Int64, Int64, Ptr{Int32}, Ptr{Int32}, Ptr{Int32}, Ptr{Int32}, Ptr{Int32}),
        h.h, L, M, nbr, deg, entry, maxl, C_NULL))
    h.graphs = true
    return h
end

# ef of the children's traversal policy (hnsw/query.jl:176-188)
hnsw_default_ef(index::MultiLevelIndex) = index.root.indices[1].traversal_policy.ef_search

function hnsw_search(index::MultiLevelIndex, data::StridedMatrix{Float32},
                     Q::StridedMatrix{Float32}, k::Int, ef::Int, probe_order::Bool)
    nq = size(Q, 2)
    ids = zeros(Int64, k, nq)
    dist = fill(Inf32, k, nq)
    Qc = dense_cols(Q)
    h = ivf_handle(index, data, Euclidean())
    lock(LOCK) do
        h.graphs || install_graphs!(h, index, size(data, 2))
    end
    GC.@preserve Qc ids dist check(ccall((:mann_ivf_hnsw_search, LIB), Cint,
        (Ptr{Cvoid}, Ptr{Float32}, Int64, Int64, Int64, Int64, Int64, Int64, Ptr{Float32}, Cint,
         Cint, Ptr{Int64}, Ptr{Float32}),
        h.h, Qc, size(Qc, 1), nq, stride(Qc, 2), k, routing_k(index).k, ef, C_NULL,
        merge_code(index.merge_strategy), Cint(probe_order), ids, dist))
    return to_neighbors(ids, dist)
end

function query(index::HIPMultiLevelHNSW, data::StridedMatrix{Float32},
               Q::StridedMatrix{Float32}, k::Integer; ef_search::Union{Nothing,Int} = nothing)
    size(Q, 1) == size(data, 1) || throw(DimensionMismatch(
        "Expected queries with $(size(data, 1)) rows, got $(size(Q, 1))"))
    size(Q, 2) == 0 && return Vector{Vector{Neighbor{Float32}}}()
    k <= 0 && return [Neighbor{Float32}[] for _ in 1:size(Q, 2)]
    ef = ef_search === nothing ? hnsw_default_ef(index) : ef_search
    return hnsw_search(index, data, Q, Int(k), ef, false)
end

function query(index::HIPMultiLevelHNSW, data::StridedMatrix{Float32},
               q::StridedVector{Float32}, k::Integer; ef_search::Union{Nothing,Int} = nothing)
    length(q) == size(data, 1) || throw(DimensionMismatch(
        "Expected query with $(size(data, 1)) rows, got $(length(q))"))
    k <= 0 && return Neighbor{Float32}[]
    ef = ef_search === nothing ? hnsw_default_ef(index) : ef_search
    return hnsw_search(index, data, reshape(dense_vec(q), :, 1), Int(k), ef, true)[1]
end

end # module
