// Continuous-batching request server (device-resident slot pool).
#pragma once

namespace phx {
class Server {
public:
    virtual ~Server() = default;
};
}  // namespace phx
