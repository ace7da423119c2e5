# Quac gates and circuits on a HIPChain.  Qrochet turns a Quac gate into a `Dense`
# on [Site.(lanes)..., Site.(lanes; dual = true)...] (ext/QrochetQuacExt.jl:7-13)
# and forwards `evolve!(::Ansatz, ::Gate)` to it (ext/QrochetQuacExt.jl:15); the
# same forwarding here reaches the device method.  A whole circuit is scheduled
# into layers of disjoint lanes, each layer one batched call.
module QrochetHIPQuacExt

using Qrochet
using QrochetHIP
using Quac: Gate, Circuit, lanes

Qrochet.evolve!(h::QrochetHIP.HIPChain, gate::Gate; kwargs...) = Qrochet.evolve!(h, Qrochet.Dense(gate); kwargs...)

function Qrochet.evolve!(h::QrochetHIP.HIPChain, circuit::Circuit; kwargs...)
    for layer in QrochetHIP.layers(collect(circuit), g -> collect(lanes(g)))
        QrochetHIP.evolve_layer!(h, [Qrochet.Dense(g) for g in layer]; kwargs...)
    end
    h
end

end
