# Stl2SdfMI355X.jl -- Julia binding of libsdf_mi355x.so (include/sdf_mi355x.h).
#
# Drop-in methods for the reference package jezekon/stl2sdf.jl: the same
# function names and argument meaning as
#   evalDistancesOnTriMesh   src/SignedDistances/dfOnTriangularMesh.jl:19
#   raycast_sign_detection   src/SignedDistances/RaycastingSignDetection.jl:337
#   remove_sdf_artifacts!    src/SignedDistances/SdfArtifactRemoval.jl:134
#   analyze_sdf_components   src/SignedDistances/SdfArtifactRemoval.jl:259
#   RBFs_smoothing           src/SdfSmoothing/RBFs4Smoothing.jl:275
#   import_stl               src/DataImport/ImportStlData.jl:10
# plus device-memory forms of stl_to_sdf's steps 4-7 (src/stl_to_sdf.jl:101-126)
# and of RBFs_smoothing, on device pointers (e.g. pointer(::ROCArray)),
# taking the reference's own TriangularMesh / Grid objects.  Julia arrays are
# passed by pointer (column-major 3 x n Float64, 1-based Int64 IEN), so no
# conversion happens on the Julia side.
module Stl2SdfMI355X

using stl2sdf.MeshGrid: TriangularMesh, Grid

const LIB = get(ENV, "SDF_MI355X_LIB", joinpath(@__DIR__, "..", "lib", "libsdf_mi355x.so"))

const _CTX = Ref{Ptr{Cvoid}}(C_NULL)

function _check(st::Int32)
    st == 0 && return nothing
    msg = unsafe_string(ccall((:sdf_last_error, LIB), Cstring, ()))
    error("libsdf_mi355x: status $st: $msg")
end

function ctx()
    if _CTX[] == C_NULL
        h = Ref{Ptr{Cvoid}}(C_NULL)
        _check(ccall((:sdf_ctx_create, LIB), Int32, (Int32, Ref{Ptr{Cvoid}}), 0, h))
        _CTX[] = h[]
    end
    return _CTX[]
end

function evalDistancesOnTriMesh(mesh::TriangularMesh, grid::Grid, points::Matrix{Float64})
    n = size(points, 2)
    dist = Vector{Float64}(undef, n)
    xp = Matrix{Float64}(undef, 3, n)
    GC.@preserve mesh grid points dist xp begin
        _check(ccall((:sdf_eval_distances_xp, LIB), Int32,
            (Ptr{Cvoid}, Ptr{Float64}, Int64, Ptr{Int64}, Int64, Ptr{Float64}, Float64,
             Ptr{Int64}, Ptr{Float64}, Int64, Ptr{Float64}, Ptr{Float64}),
            ctx(), mesh.X, mesh.nnp, mesh.IEN, mesh.nel, grid.AABB_min, grid.cell_size,
            grid.N, points, n, dist, xp))
    end
    return dist, xp
end

function raycast_sign_detection(mesh::TriangularMesh, grid::Grid, points::Matrix{Float64};
                                min_rays::Int = 6, max_rays::Int = 30,
                                confidence_threshold::Float64 = 0.6,
                                use_winding_fallback::Bool = true)
    n = size(points, 2)
    signs = Vector{Float64}(undef, n)
    conf = Vector{Float64}(undef, n)
    GC.@preserve mesh grid points signs conf begin
        _check(ccall((:sdf_raycast_signs, LIB), Int32,
            (Ptr{Cvoid}, Ptr{Float64}, Int64, Ptr{Int64}, Int64, Ptr{Float64}, Float64,
             Ptr{Int64}, Ptr{Float64}, Int64, Int32, Int32, Float64, Int32,
             Ptr{Float64}, Ptr{Float64}),
            ctx(), mesh.X, mesh.nnp, mesh.IEN, mesh.nel, grid.AABB_min, grid.cell_size,
            grid.N, points, n, min_rays, max_rays, confidence_threshold,
            Int32(use_winding_fallback), signs, conf))
    end
    return signs, conf
end

function remove_sdf_artifacts!(sdf_values::Vector{Float64}, grid::Grid;
                               threshold::Float64 = 0.0, min_component_ratio::Float64 = 0.01)
    length(sdf_values) == grid.ngp ||
        error("SDF values length ($(length(sdf_values))) doesn't match grid points ($(grid.ngp))")
    n = Ref{Int64}(0)
    GC.@preserve sdf_values grid begin
        _check(ccall((:sdf_remove_artifacts, LIB), Int32,
            (Ptr{Cvoid}, Ptr{Float64}, Ptr{Int64}, Float64, Float64, Ref{Int64}),
            ctx(), sdf_values, grid.N, threshold, min_component_ratio, n))
    end
    return Int(n[])
end

function analyze_sdf_components(sdf_values::Vector{Float64}, grid::Grid;
                                threshold::Float64 = 0.0)
    nc = Ref{Int64}(0)
    GC.@preserve sdf_values grid begin
        _check(ccall((:sdf_analyze_components, LIB), Int32,
            (Ptr{Cvoid}, Ptr{Float64}, Ptr{Int64}, Float64, Ptr{Int64}, Ptr{Int64},
             Ptr{Int64}, Int64, Ref{Int64}),
            ctx(), sdf_values, grid.N, threshold, C_NULL, C_NULL, C_NULL, 0, nc))
        roots = Vector{Int64}(undef, nc[])
        sizes = Vector{Int64}(undef, nc[])
        _check(ccall((:sdf_analyze_components, LIB), Int32,
            (Ptr{Cvoid}, Ptr{Float64}, Ptr{Int64}, Float64, Ptr{Int64}, Ptr{Int64},
             Ptr{Int64}, Int64, Ref{Int64}),
            ctx(), sdf_values, grid.N, threshold, C_NULL, roots, sizes, nc[], nc))
    end
    return Dict{Int,Int}(r + 1 => s for (r, s) in zip(roots, sizes))
end

function RBFs_smoothing(dist::Vector, my_grid::Grid, Is_interpolation::Bool, smooth::Int,
                        threshold::Float64 = 1e-3)
    d = Vector{Float64}(dist)
    nf = prod(my_grid.N .* smooth .+ 1)
    fine = Vector{Float32}(undef, nf)
    GC.@preserve d my_grid fine begin
        _check(ccall((:sdf_rbf_smoothing, LIB), Int32,
            (Ptr{Cvoid}, Ptr{Float64}, Ptr{Int64}, Float64, Int32, Int32, Float64, Float64,
             Ptr{Float32}, Ptr{Int32}, Ptr{Float64}),
            ctx(), d, my_grid.N, my_grid.cell_size, Int32(Is_interpolation), smooth,
            threshold, 0.0, fine, C_NULL, C_NULL))
    end
    new_grid = Grid(my_grid.AABB_min, my_grid.AABB_max, maximum(my_grid.N) * smooth, 0)
    return fine, new_grid
end

# import_stl(filename) -> (X, IEN): the package's STL readers and its 1e-10
# vertex merge (ImportStlData.jl:102-309), in the library's host code.
function import_stl(filename::String)
    nnp = Ref{Int64}(0)
    nel = Ref{Int64}(0)
    _check(ccall((:sdf_import_stl, LIB), Int32,
        (Cstring, Ptr{Float64}, Ref{Int64}, Ptr{Int64}, Ref{Int64}),
        filename, C_NULL, nnp, C_NULL, nel))
    X = Matrix{Float64}(undef, 3, nnp[])
    IEN = Matrix{Int64}(undef, 3, nel[])
    _check(ccall((:sdf_import_stl, LIB), Int32,
        (Cstring, Ptr{Float64}, Ref{Int64}, Ptr{Int64}, Ref{Int64}),
        filename, X, nnp, IEN, nel))
    return X, IEN
end

# stl_to_sdf.jl:101-126 into device memory: distances, signs, their product
# and (remove_artifacts) remove_sdf_artifacts! on the device; d_sdf points to
# grid.ngp Float64 in device memory.  Returns the number of flipped nodes.
function signed_fused!(d_sdf::Ptr{Float64}, mesh::TriangularMesh, grid::Grid;
                       remove_artifacts::Bool = false, artifact_ratio::Float64 = 0.01)
    nflipped = Ref{Int64}(0)
    GC.@preserve mesh grid begin
        _check(ccall((:sdf_mesh_set, LIB), Int32,
            (Ptr{Cvoid}, Ptr{Float64}, Int64, Ptr{Int64}, Int64),
            ctx(), mesh.X, mesh.nnp, mesh.IEN, mesh.nel))
        _check(ccall((:sdf_signed_fused, LIB), Int32,
            (Ptr{Cvoid}, Ptr{Float64}, Float64, Ptr{Int64}, Int32, Float64, Ptr{Float64},
             Ref{Int64}),
            ctx(), grid.AABB_min, grid.cell_size, grid.N, Int32(remove_artifacts),
            artifact_ratio, d_sdf, nflipped))
    end
    return nflipped[]
end

# RBFs_smoothing on device memory: d_dist holds my_grid.ngp Float64, d_fine
# receives prod(N .* smooth .+ 1) Float32.  Returns the fine grid.
function RBFs_smoothing!(d_fine::Ptr{Float32}, d_dist::Ptr{Float64}, my_grid::Grid,
                         Is_interpolation::Bool, smooth::Int, threshold::Float64 = 1e-3)
    GC.@preserve my_grid begin
        _check(ccall((:sdf_rbf_smoothing_device, LIB), Int32,
            (Ptr{Cvoid}, Ptr{Float64}, Ptr{Int64}, Float64, Int32, Int32, Float64, Float64,
             Ptr{Float32}, Ptr{Int32}, Ptr{Float64}),
            ctx(), d_dist, my_grid.N, my_grid.cell_size, Int32(Is_interpolation), smooth,
            threshold, 0.0, d_fine, C_NULL, C_NULL))
    end
    return Grid(my_grid.AABB_min, my_grid.AABB_max, maximum(my_grid.N) * smooth, 0)
end

# Learned search radius (no counterpart in the reference): fit once per mesh
# and grid, then compute the same field as signed_distance.
function learned_fit(mesh::TriangularMesh, grid::Grid; n_coarse::Int = 128,
                     steps::Int = 1500, batch::Int = 4096, lr::Float32 = 2f-3,
                     seed::UInt64 = UInt64(1), quantile::Float64 = 0.99)
    stats = zeros(Float64, 9)
    GC.@preserve mesh grid stats begin
        _check(ccall((:sdf_mesh_set, LIB), Int32,
            (Ptr{Cvoid}, Ptr{Float64}, Int64, Ptr{Int64}, Int64),
            ctx(), mesh.X, mesh.nnp, mesh.IEN, mesh.nel))
        _check(ccall((:sdf_learned_fit, LIB), Int32,
            (Ptr{Cvoid}, Ptr{Float64}, Float64, Ptr{Int64}, Int32, Int32, Int32, Float32,
             UInt64, Float64, Ptr{Float64}, Int32),
            ctx(), grid.AABB_min, grid.cell_size, grid.N, n_coarse, steps, batch, lr,
            seed, quantile, stats, Int32(9)))
    end
    return stats
end

function learned_signed_distance(mesh::TriangularMesh, grid::Grid)
    out = Vector{Float64}(undef, grid.ngp)
    GC.@preserve mesh grid out begin
        _check(ccall((:sdf_learned_signed_distance, LIB), Int32,
            (Ptr{Cvoid}, Ptr{Float64}, Int64, Ptr{Int64}, Int64, Ptr{Float64}, Float64,
             Ptr{Int64}, Ptr{Float64}),
            ctx(), mesh.X, mesh.nnp, mesh.IEN, mesh.nel, grid.AABB_min, grid.cell_size,
            grid.N, out))
    end
    return out
end

end # module
