// Cache-blocked multi-gate passes: shared host/device definitions.
//
// A pass streams a state through registers once, tile by tile, and applies
// a run of one-qubit gates (with any controls) to every tile on the way.  A
// tile is 2^12 amplitudes: index bits 0..5 (the lane bits: 64 consecutive
// amplitudes per wave, so every load and store is one contiguous 1 KiB run
// per wave instruction) and six more index bits chosen per pass (the high
// slot bits).  The remaining index bits are fixed per tile.
//
// A workgroup holds its tile in registers: 4 waves x 64 lanes x 16
// amplitudes.  Of the six high slot bits, four are register bits (the 16
// amplitudes of a thread differ in them) and two are wave bits; which is
// which is the pass's current mapping.  An exchange op moves the tile
// through LDS into another mapping; one register bit keeps its slot across
// it, so it runs as two rounds over half the tile each (a half-tile LDS
// buffer per workgroup).  Gates act
//   - on a register bit: pairs of a thread's own registers;
//   - on a lane bit: the partner amplitude comes from lane ^ 2^b;
//   - diagonally on any qubit: a per-amplitude factor, no partner at all.
//
// The index bits all of a thread's registers share (tile base, lane, wave
// bits) form its thread word; a control on one of those qubits is a test of
// the thread word, a control on a register bit a test of the register
// number.  The planner splits each gate's masks that way under the mapping
// current at the op.
#pragma once

#include <cstdint>

namespace dkp {

constexpr int LANE_BITS = 6;
constexpr int HIGH_BITS = 6;
constexpr int PASS_TILE_BITS = LANE_BITS + HIGH_BITS; // 12
constexpr int REG_BITS = 4;
constexpr int WAVE_BITS = HIGH_BITS - REG_BITS;       // 2
constexpr int REGS = 1 << REG_BITS;                   // 16 amplitudes per thread
constexpr int THREADS = 64 << WAVE_BITS;              // 256

enum OpKind : uint32_t {
	OP_REG = 0,         // dense 2x2 on register bit a
	OP_LANE = 1,        // dense 2x2 on lane bit (qubit) a
	OP_DIAG_REG = 2,    // diagonal on register bit a
	OP_DIAG_THREAD = 3, // diagonal on qubit a, a bit of the thread word
	OP_EXCH = 4,        // new mapping: a = 4 x 3-bit register slots, b = 2 x
	                    // 3-bit wave slots (high slot bit indices)
	OP_REG_MUX = 5,     // dense 2x2 on register bit a, matrix chosen per
	                    // amplitude by a selector qubit (b): m if it is 0,
	                    // the next record's m if it is 1
	OP_LANE_MUX = 6,    // the same on lane bit (qubit) a
	OP_MDATA = 7,       // second matrix of the preceding mux op
};

// Selector of a mux op (field b): SEL_REG | j for register bit j (register
// targets only), or the qubit of a thread-word bit >= 6, the same for the
// whole wave.
constexpr uint32_t SEL_REG = 0x100;

// One op of a pass (96 bytes).  rmask: register-bit controls in bits 0..3,
// register-bit anticontrols in bits 8..11; tcm / tam: control and
// anticontrol qubits of the thread word.  Mux ops carry no controls.
struct alignas(16) PassOp {
	uint32_t kind, a, b, rmask;
	uint64_t tcm, tam;
	double m[8]; // 2x2 row-major (re, im); diagonal ops use m[0..1], m[6..7]
};

// Per-pass launch description.
struct PassDesc {
	uint8_t hpos[HIGH_BITS];    // index bit of high slot bit h
	uint8_t hsorted[HIGH_BITS]; // the same bits, ascending
	uint32_t nops;
	uint32_t finite;            // every op's matrix entries are finite
	// zwrite, zmask (set at launch): index bits known to be zero in the
	// input; a tile whose fixed bits meet zmask holds only zeros, and so
	// does its output, which is written as zeros (zwrite) or, where the
	// output already holds those zeros (in place), not written at all,
	// without reading the tile or running the ops
	uint32_t zwrite;
	uint64_t zmask;
	// kmask: index bits known to be zero in the input where the input
	// memory need not hold those zeros (the batched engine's pool is not
	// cleared): an amplitude whose index meets kmask is taken as zero
	// instead of read.  0 reads every amplitude of a tile it runs
	uint64_t kmask;
};

// One register's pass within a round of batched passes: its tiles are the
// round's items first .. first + 2^(nq - 12) - 1, its state starts at
// amplitude `off` of the pool, its ops at op_begin of the round's op array.
struct PassEntry {
	uint64_t first;
	uint64_t off;
	PassDesc pd;
	uint32_t op_begin;
	uint32_t nq;
};

} // namespace dkp
