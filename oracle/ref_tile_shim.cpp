// C-ABI entry to the reference's own GPU-RFOR encoder, rleBinPack
// (tile_based/src/rlebinpack.cpp:9-146), compiled where it lies with its
// main renamed.  Test infrastructure only (oracle/_ref).
#include <cstdint>
#include <utility>

using uint = unsigned int;
std::pair<uint, uint> rleBinPack(uint*& in, uint*& value, uint*& run_length, uint*& val_offsets, uint*& rl_offsets,
                                 uint num_entries);

extern "C" void ref_rle_bin_pack(uint32_t* in, uint32_t n, uint32_t* value, uint32_t* run_length,
                                 uint32_t* val_offsets, uint32_t* rl_offsets, uint64_t* n_val, uint64_t* n_rl) {
    uint* pin = in;
    uint* pv = value;
    uint* pr = run_length;
    uint* pvo = val_offsets;
    uint* pro = rl_offsets;
    const std::pair<uint, uint> r = rleBinPack(pin, pv, pr, pvo, pro, n);
    *n_val = r.first;
    *n_rl = r.second;
}
