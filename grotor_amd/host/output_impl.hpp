// .out formatting: src/output.hpp:79-112 and src/output_writer.hpp:29-43.
// Per design net: the name, "(", a via line from metal 0 to metal 1 for every
// pin on metal 0, the extra_routes pairs, every subnet's segments shifted up
// by one metal layer (low endpoint first), then ")".
#pragma once

#include <charconv>
#include <string>

namespace grotor {

template <class RoutesOf>
std::string format_output(const Design& design, RoutesOf routes_of) {
    const int X = design.grid.X, Y = design.grid.Y, XY = X * Y;
    std::string out;
    out.reserve(design.nets.size() * 64);
    char buffer[96];
    auto segment = [&](int p0, int p1) {
        if (p0 > p1) std::swap(p0, p1);
        const int values[6] = {p0 / Y % X, p0 % Y, p0 / XY, p1 / Y % X, p1 % Y, p1 / XY};
        char* pos = buffer;
        for (int i = 0; i < 6; ++i) {
            pos = std::to_chars(pos, buffer + sizeof(buffer), values[i]).ptr;
            *pos++ = i == 5 ? '\n' : ' ';
        }
        out.append(buffer, pos);
    };
    for (const auto& net : design.nets) {
        out += net.name;
        out += "\n(\n";
        for (int pin : net.pins)
            if (pin < XY) segment(pin, pin + XY);
        for (size_t i = 0; i + 1 < net.extra_routes.size(); i += 2) segment(net.extra_routes[i], net.extra_routes[i + 1]);
        for (int id : net.subnets) {
            const int* r = routes_of(id);
            for (int j = 1; j < r[0]; j += 2) segment(r[j] + XY, r[j + 1] + XY);
        }
        out += ")\n";
    }
    return out;
}

}  // namespace grotor
