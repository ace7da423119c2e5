# Yao circuits on a HIPChain.  Qrochet turns a Yao circuit into a `Quantum` network
# (ext/QrochetYaoExt.jl:15-48): nested chains are flattened (:7-13), a `Swap`
# exchanges the wires of its two lanes without adding a tensor (:25-29), every other
# block becomes a tensor on the wires at its locations (:31-39), and lane l's input
# and output sites are the two ends of the wire that ends at lane l (:44-47).  Here a
# wire is identified with the MPS site where it ends, so each block becomes a
# `Dense` gate (outputs first, like ext/QrochetQuacExt.jl:7-13) on the sites where
# its wires end, no SWAP is applied to the state, and the gates run layer by layer
# (one batched call per layer of two-site gates).
module QrochetHIPYaoExt

using Qrochet
using Qrochet: Dense, Operator, Site, outputs, id
using QrochetHIP
using Yao

flatten_blocks(x) = any(b -> b isa ChainBlock, subblocks(x)) ? flatten_blocks(Yao.Optimise.eliminate_nested(x)) : x

function yao_gates(circuit::AbstractBlock)
    @assert nlevel(circuit) == 2 "Only support 2-level qubits"
    n = nqubits(circuit)
    blocks = collect(flatten_blocks(circuit))
    lane_wire = collect(1:n)
    for g in blocks
        if g isa Swap
            (a, b) = occupied_locs(g)
            lane_wire[a], lane_wire[b] = lane_wire[b], lane_wire[a]
        end
    end
    end_slot = zeros(Int, n)
    for (lane, w) in enumerate(lane_wire)
        end_slot[w] = lane
    end
    lane_wire = collect(1:n)
    gates = Dense[]
    for g in blocks
        if g isa Swap
            (a, b) = occupied_locs(g)
            lane_wire[a], lane_wire[b] = lane_wire[b], lane_wire[a]
            continue
        end
        sites = [end_slot[lane_wire[l]] for l in occupied_locs(g)]
        k = length(sites)
        array = reshape(Matrix{ComplexF64}(mat(content(g))), fill(2, 2k)...)
        push!(gates, Dense(Operator(), array; sites = Site[Site.(sites)..., Site.(sites; dual = true)...]))
    end
    gates
end

function Qrochet.evolve!(h::QrochetHIP.HIPChain, circuit::AbstractBlock; kwargs...)
    for layer in QrochetHIP.layers(yao_gates(circuit), g -> map(id, outputs(g)))
        QrochetHIP.evolve_layer!(h, layer; kwargs...)
    end
    h
end

end
