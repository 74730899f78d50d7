// particleSamples -- the part of the reference's event container that the
// HBT path uses (src/particleSamples.h): read_in_mode 10 (gzip text,
// particleSamples.cpp:1247-1286) and 9 (binary records, :1203-1245) in
// buffers of whole events, species filter
// (a Monte-Carlo id, 99999 for all, or a species class), optional separate
// mixed-event file
// (read_in_real_mixed_events), the same getters.  Particles are kept as
// structure-of-arrays so that HBT_correlation can hand them to the GPU
// engine without repacking.
#ifndef HBT_AMD_PARTICLESAMPLES_H
#define HBT_AMD_PARTICLESAMPLES_H

#include <memory>
#include <string>
#include <vector>

#include "ParameterReader.h"
#include "Random.h"
#include "hbt_hip.h"
#include "particle_info.h"

class particleSamples {
  public:
    particleSamples(ParameterReader &paraRdr, std::string path,
                    std::shared_ptr<RandomUtil::Random> ran_gen);
    ~particleSamples();

    void read_in_particle_samples_and_filter();
    void read_in_particle_samples_mixed_event_and_filter();
    bool end_of_file() const;
    bool end_of_file_mixed_event() const;

    int get_event_buffer_size() const { return event_buffer_size; }
    int get_number_of_events() const { return (int)same_.off.size() - 1; }
    int get_number_of_mixed_events() const { return (int)mixed().off.size() - 1; }
    int get_number_of_particles(int event_id) const {
        return (int)(same_.off[event_id + 1] - same_.off[event_id]);
    }
    int get_number_of_particles_mixed_event(int event_id) const {
        return (int)(mixed().off[event_id + 1] - mixed().off[event_id]);
    }
    particle_info get_particle(int event_id, int part_id) const {
        return same_.get(event_id, part_id, particle_monval);
    }
    particle_info get_particle_from_mixed_event(int event_id, int part_id) const {
        return mixed().get(event_id, part_id, particle_monval);
    }

    // Structure-of-arrays view of the current buffers (for the GPU engine):
    // soa = {px, py, pz, E, x, y, z, t}, offsets of nev+1 events.
    struct Events {
        std::vector<int64_t> off{0};
        std::vector<double> cols[8];
        std::vector<double> mass;
        std::vector<int32_t> monval;
        particle_info get(int ev, int i, int monval) const;
    };
    const Events &same_events() const { return same_; }
    const Events &mixed_events() const { return mixed(); }
    bool mixed_is_separate() const { return read_mixed_events; }

  private:
    const ParameterReader paraRdr_;
    std::string path_;
    int event_buffer_size;
    int particle_monval;
    bool read_mixed_events;
    hbt_reader *reader = nullptr, *reader_mixed = nullptr;
    Events same_, mixed_;
    const Events &mixed() const { return read_mixed_events ? mixed_ : same_; }
    static void read(hbt_reader *rd, Events &ev);
};

#endif
