// netgen.h -- the host-side network object behind nh_net: the flat
// simulator arrays (the fields of nh_net_desc) plus, for networks read from
// .ne2 files, the editor-level tables callers address by index.
#pragma once
#include <stdint.h>

#include <string>
#include <utility>
#include <vector>

struct nh_net {
    uint32_t n = 0;
    std::vector<uint32_t> src_ptr, tgt;
    std::vector<float> w;
    std::vector<uint16_t> delay;
    std::vector<float> tau;  // per synapse (empty: every tau is 0)
    std::vector<uint8_t> model_ids, model;
    uint32_t n_params = 4, n_state = 2;
    std::vector<float> p, x0;
    uint32_t n_inputs = 0;
    std::vector<uint32_t> stim_ptr, stim_input;
    std::vector<float> stim_w;
    // .ne2 networks only
    std::vector<std::string> type_names;
    std::vector<uint16_t> neuron_type;
    std::vector<std::string> input_names;
    std::vector<std::pair<std::string, std::vector<uint32_t>>> groups;
};
