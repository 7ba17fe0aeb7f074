/*
 * Writer factory and base class
 * (reference: mbsolve-lib/include/mbsolve/lib/writer.hpp,
 * mbsolve-lib/src/writer.cpp).
 */
#ifndef MBSOLVE_LIB_WRITER_H
#define MBSOLVE_LIB_WRITER_H

#include <functional>
#include <map>
#include <memory>
#include <string>
#include <vector>
#include <mbsolve/lib/device.hpp>
#include <mbsolve/lib/result.hpp>
#include <mbsolve/lib/scenario.hpp>

namespace mbsolve {

class writer
{
public:
    typedef std::function<std::shared_ptr<writer>()> bootstrap_t;

private:
    static std::map<std::string, bootstrap_t>& bootstraps();

protected:
    writer(const std::string& name, const std::string& extension);

    std::string m_name;
    std::string m_ext;

public:
    virtual ~writer();

    static std::vector<std::string> get_avail_writers();

    static void register_bootstrap(const std::string& name, bootstrap_t b);

    template<typename T>
    static void register_writer(const std::string& name)
    {
        register_bootstrap(name, []() { return std::make_shared<T>(); });
    }

    static std::shared_ptr<writer> create_instance(const std::string& name);

    const std::string& get_name() const;

    const std::string& get_extension() const;

    virtual void write(
        const std::string& file,
        const std::vector<std::shared_ptr<result> >& results,
        std::shared_ptr<const device> dev,
        std::shared_ptr<const scenario> scen) const = 0;
};
}

#endif
