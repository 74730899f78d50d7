// particleSamples.cpp -- see particleSamples.h.  Reader modes 10 and 9 of
// the reference (particleSamples.cpp:20-221 for the keys it reads and the
// file names :131-152, :480-552 for the buffer calls, :1203-1346 for the
// formats and the species filter).
#include "particleSamples.h"

#include <cmath>
#include <cstdlib>
#include <iostream>
#include <sstream>

particleSamples::particleSamples(ParameterReader &paraRdr, std::string path,
                                 std::shared_ptr<RandomUtil::Random> /*ran_gen*/)
    : paraRdr_(paraRdr), path_(path) {
    event_buffer_size = (int)paraRdr_.getVal("event_buffer_size");
    const int read_in_mode = (int)paraRdr_.getVal("read_in_mode");
    particle_monval = (int)paraRdr_.getVal("particle_monval");
    read_mixed_events = paraRdr_.getVal("read_in_real_mixed_events") == 1;
    if (read_in_mode != 10 && read_in_mode != 9) {
        std::cout << "particleSamples:: Error: read_in_mode " << read_in_mode
                  << " is not supported (modes 10, gzip text, and 9, binary, are)"
                  << std::endl;
        exit(1);
    }
    const bool bin = read_in_mode == 9;
    const std::string fn = path_ + (bin ? "/particle_list.bin" : "/particle_samples.gz");
    const std::string fn_mixed =
        path_ + (bin ? "/particle_list_mixed_event.bin" : "/particle_samples_mixed_event.gz");
    if (hbt_reader_open_mode(fn.c_str(), read_in_mode, event_buffer_size, particle_monval,
                             &reader) != HBT_OK) {
        std::cout << "particleSamples:: Error: input file: " << fn << " can not open!"
                  << std::endl;
        exit(1);
    }
    if (read_mixed_events &&
        hbt_reader_open_mode(fn_mixed.c_str(), read_in_mode, event_buffer_size, particle_monval,
                             &reader_mixed) != HBT_OK) {
        std::cout << "particleSamples:: Error: input file: " << fn_mixed << " can not open!"
                  << std::endl;
        exit(1);
    }
    // species classes pick by charge, baryon number or strangeness from the
    // PDG table the reference's particle_decay reads (particle_decay.cpp:
    // 26-36): EOS/pdg.dat, or EOS/pdg_weak.dat with weak feed-down
    const int m = particle_monval;
    if (m == 9999 || m == 9998 || m == -9998 || m == 9997 || m == -9997 || m == 9996 ||
        m == -9996) {
        const std::string pdg = paraRdr_.getVal("resonance_weak_feed_down_flag", 0) == 1
                                    ? "EOS/pdg_weak.dat"
                                    : "EOS/pdg.dat";
        hbt_species_table *t = nullptr;
        if (hbt_species_read(pdg.c_str(), &t) != HBT_OK) {
            std::cout << "can not find the file " << pdg << "!" << std::endl;
            exit(1);
        }
        hbt_reader_set_species(reader, t);
        if (reader_mixed) hbt_reader_set_species(reader_mixed, t);
        hbt_species_free(t);
    }
}

particleSamples::~particleSamples() {
    hbt_reader_close(reader);
    hbt_reader_close(reader_mixed);
}

void particleSamples::read(hbt_reader *rd, Events &ev) {
    const int nev = hbt_reader_next(rd);
    if (nev < 0) {
        // the reference's particle_decay lookup throws for a monval its
        // table does not hold
        std::cout << "particleSamples:: Error: particle " << hbt_reader_bad_monval(rd)
                  << " is not in the PDG table" << std::endl;
        exit(1);
    }
    const int64_t *off = nullptr;
    const double *soa[8] = {};
    const double *mass = nullptr;
    hbt_reader_buffer(rd, &off, soa, &mass, nullptr);
    ev.off.assign(off, off + nev + 1);
    const int64_t n = off[nev];
    for (int a = 0; a < 8; a++) ev.cols[a].assign(soa[a], soa[a] + n);
    ev.mass.assign(mass, mass + n);
    const int32_t *mon = nullptr;
    hbt_reader_monval(rd, &mon);
    ev.monval.assign(mon, mon + n);
}

void particleSamples::read_in_particle_samples_and_filter() { read(reader, same_); }

void particleSamples::read_in_particle_samples_mixed_event_and_filter() {
    // without read_in_real_mixed_events the mixed list is the same-event
    // list (particleSamples.cpp:528-531)
    if (read_mixed_events) read(reader_mixed, mixed_);
}

bool particleSamples::end_of_file() const { return hbt_reader_eof(reader) != 0; }
bool particleSamples::end_of_file_mixed_event() const {
    return read_mixed_events ? hbt_reader_eof(reader_mixed) != 0 : end_of_file();
}

particle_info particleSamples::Events::get(int ev, int i, int monval) const {
    const int64_t k = off[ev] + i;
    particle_info p;
    p.monval = monval == 99999 ? 0 : monval;
    p.mass = mass[k];
    p.baryon = 0;
    p.px = cols[0][k];
    p.py = cols[1][k];
    p.pz = cols[2][k];
    p.E = cols[3][k];
    p.x = cols[4][k];
    p.y = cols[5][k];
    p.z = cols[6][k];
    p.t = cols[7][k];
    // as particleSamples::boostParticles fills them (rapidity shift 0)
    p.pT = sqrt(p.px * p.px + p.py * p.py);
    const double mT = sqrt(p.pT * p.pT + p.mass * p.mass);
    p.phi_p = atan2(p.py, p.px);
    p.rap_y = asinh(p.pz / mT);
    p.rap_eta = asinh(p.pz / p.pT);
    return p;
}
