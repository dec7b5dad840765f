// bindings.cpp -- Python module `_app`: the Application API of the reference's module
// (src/python/bindings.cpp:28-100) over the headless host layer, plus numpy accessors for
// the displayed image and the passes' knobs.  step() and run_until_closed() release the GIL.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include "rr_host.h"

namespace py = pybind11;
using rrh::app::Application;
using namespace rrh::passes;

PYBIND11_MODULE(_app, m) {
    m.doc() = "Headless research renderer on MI355X (HIP): the reference Application API.";

    auto image_f32 = [](const rrh::rhi::Image& im, rrh::rhi::Device* d) {
        py::array_t<float> out({(py::ssize_t)im.extent().height, (py::ssize_t)im.extent().width, (py::ssize_t)4});
        im.download(*d, out.mutable_data(), (size_t)out.nbytes());
        return out;
    };
    py::class_<GBufferPass>(m, "GBufferPass")
        .def("is_enabled", &GBufferPass::is_enabled)
        .def_property_readonly("name", [](const GBufferPass& p) { return std::string(p.name()); })
        .def("position", [image_f32](const GBufferPass& p) { return image_f32(p.position_image(), p.device()); })
        .def("normal", [image_f32](const GBufferPass& p) { return image_f32(p.normal_image(), p.device()); })
        .def("material_id", [](const GBufferPass& p) {
            const auto& im = p.material_id_image();
            py::array_t<uint32_t> out({(py::ssize_t)im.extent().height, (py::ssize_t)im.extent().width});
            im.download(*p.device(), out.mutable_data(), (size_t)out.nbytes());
            return out;
        });
    py::class_<TonemapPass>(m, "TonemapPass")
        .def_readwrite("exposure", &TonemapPass::exposure)
        .def("is_enabled", &TonemapPass::is_enabled)
        .def_property_readonly("name", [](const TonemapPass& p) { return std::string(p.name()); });
    py::class_<PathTracerPass>(m, "PathTracerPass")
        .def_readwrite("max_bounces", &PathTracerPass::max_bounces)
        .def_readwrite("rr_start", &PathTracerPass::rr_start)
        .def("is_enabled", &PathTracerPass::is_enabled)
        .def_property_readonly("name", [](const PathTracerPass& p) { return std::string(p.name()); });
    py::class_<AccumulatePass>(m, "AccumulatePass")
        .def_readwrite("fp64_enable", &AccumulatePass::fp64_enable)
        .def_readonly("accumulated_spp", &AccumulatePass::accumulated_spp)
        .def("is_enabled", &AccumulatePass::is_enabled);
    py::class_<ReSTIRDIPass>(m, "ReSTIRDIPass")
        .def_readwrite("temporal_enable", &ReSTIRDIPass::temporal_enable)
        .def_readwrite("spatial_enable", &ReSTIRDIPass::spatial_enable)
        .def_readwrite("num_initial_candidates", &ReSTIRDIPass::num_initial_candidates)
        .def_readwrite("spatial_num_neighbors", &ReSTIRDIPass::spatial_num_neighbors)
        .def_readwrite("spatial_radius", &ReSTIRDIPass::spatial_radius)
        .def("reset_history", &ReSTIRDIPass::reset_history)
        .def("is_enabled", &ReSTIRDIPass::is_enabled);
    py::class_<ReSTIRPTPass>(m, "ReSTIRPTPass")
        .def_readwrite("include_direct_lighting", &ReSTIRPTPass::include_direct_lighting)
        .def_readwrite("temporal_enable", &ReSTIRPTPass::temporal_enable)
        .def_readwrite("spatial_enable", &ReSTIRPTPass::spatial_enable)
        .def_readwrite("indirect_strength", &ReSTIRPTPass::indirect_strength)
        .def_readwrite("ray_bias", &ReSTIRPTPass::ray_bias)
        .def_readwrite("num_initial_candidates", &ReSTIRPTPass::num_initial_candidates)
        .def_readwrite("max_bounces", &ReSTIRPTPass::max_bounces)
        .def_readwrite("spatial_num_neighbors", &ReSTIRPTPass::spatial_num_neighbors)
        .def_readwrite("spatial_radius", &ReSTIRPTPass::spatial_radius)
        .def_readwrite("guiding_enable", &ReSTIRPTPass::guiding_enable)
        .def_readwrite("guiding_alpha", &ReSTIRPTPass::guiding_alpha)
        .def_readwrite("grid_base_cell_size", &ReSTIRPTPass::grid_base_cell_size)
        .def_readwrite("grid_distance_slope", &ReSTIRPTPass::grid_distance_slope)
        .def_readwrite("grid_min_samples", &ReSTIRPTPass::grid_min_samples)
        .def_readwrite("grid_kappa_max", &ReSTIRPTPass::grid_kappa_max)
        .def("reset_history", &ReSTIRPTPass::reset_history)
        .def("is_enabled", &ReSTIRPTPass::is_enabled)
        .def_property_readonly("name", [](const ReSTIRPTPass& p) { return std::string(p.name()); });
    py::class_<ReSTIRGIPass>(m, "ReSTIRGIPass")
        .def_readwrite("include_direct_lighting", &ReSTIRGIPass::include_direct_lighting)
        .def_readwrite("temporal_enable", &ReSTIRGIPass::temporal_enable)
        .def_readwrite("spatial_enable", &ReSTIRGIPass::spatial_enable)
        .def_readwrite("indirect_strength", &ReSTIRGIPass::indirect_strength)
        .def_readwrite("ray_bias", &ReSTIRGIPass::ray_bias)
        .def_readwrite("num_initial_candidates", &ReSTIRGIPass::num_initial_candidates)
        .def_readwrite("spatial_num_neighbors", &ReSTIRGIPass::spatial_num_neighbors)
        .def_readwrite("spatial_radius", &ReSTIRGIPass::spatial_radius)
        .def("reset_history", &ReSTIRGIPass::reset_history)
        .def("is_enabled", &ReSTIRGIPass::is_enabled)
        .def_property_readonly("name", [](const ReSTIRGIPass& p) { return std::string(p.name()); });
    py::class_<AtrousPass>(m, "AtrousPass")
        .def_readwrite("iterations", &AtrousPass::iterations)
        .def_readwrite("sigma_position", &AtrousPass::sigma_position)
        .def_readwrite("sigma_normal", &AtrousPass::sigma_normal)
        .def_readwrite("sigma_luminance", &AtrousPass::sigma_luminance)
        .def_readwrite("fast_math", &AtrousPass::fast_math)
        .def("is_enabled", &AtrousPass::is_enabled);
    py::class_<GuidingPass>(m, "GuidingPass")
        .def_property(
            "mode", [](const GuidingPass& g) { return g.mode == GuidingPass::Mode::VMM ? "vmm" : "mlp"; },
            [](GuidingPass& g, const std::string& s) {
                if (s == "vmm") g.mode = GuidingPass::Mode::VMM;
                else if (s == "mlp") g.mode = GuidingPass::Mode::MLP;
                else throw std::invalid_argument("GuidingPass.mode: 'vmm' or 'mlp'");
            })
        .def_readwrite("guiding_enable", &GuidingPass::guiding_enable)
        .def_readwrite("guiding_alpha", &GuidingPass::guiding_alpha)
        .def_readwrite("max_bounces", &GuidingPass::max_bounces)
        .def_readwrite("grid_resolution", &GuidingPass::grid_resolution)
        .def_readwrite("grid_kappa_max", &GuidingPass::grid_kappa_max)
        .def_readwrite("grid_decay", &GuidingPass::grid_decay)
        .def_readwrite("record_vertices", &GuidingPass::record_vertices)
        .def_readwrite("mlp_batch", &GuidingPass::mlp_batch)
        .def_readwrite("mlp_learning_rate", &GuidingPass::mlp_learning_rate)
        .def_readwrite("mlp_steps_per_frame", &GuidingPass::mlp_steps_per_frame)
        .def_property_readonly("last_loss", &GuidingPass::last_loss)
        .def("is_enabled", &GuidingPass::is_enabled);

    py::class_<Application>(m, "Application")
        .def(py::init<>())
        // lifecycle
        .def("initialize", &Application::initialize, "Create the device and passes; the Cornell box loads at the first step.")
        .def("shutdown", &Application::shutdown)
        .def("step", &Application::step, py::call_guard<py::gil_scoped_release>(), "Render exactly one frame.")
        .def("should_close", &Application::should_close)
        .def("run_until_closed", &Application::run_until_closed, py::call_guard<py::gil_scoped_release>())
        // scene
        .def("load_cornell", &Application::load_cornell)
        .def("load_gltf", &Application::load_gltf, py::arg("path"))
        .def("load_procedural", &Application::load_procedural, py::arg("kind"), py::arg("seed") = 1,
             py::arg("scale") = 1.0f)
        .def_property_readonly("scene_name", &Application::scene_name)
        .def_property_readonly("last_scene_error", &Application::last_scene_error)
        // camera
        .def(
            "set_camera_look_at",
            [](Application& self, std::array<float, 3> eye, std::array<float, 3> center, std::array<float, 3> up) {
                self.set_camera_look_at(eye[0], eye[1], eye[2], center[0], center[1], center[2], up[0], up[1], up[2]);
            },
            py::arg("eye"), py::arg("center"), py::arg("up") = std::array<float, 3>{0.0f, 1.0f, 0.0f})
        .def_property_readonly("camera_position",
                               [](const Application& a) { auto v = a.camera_position(); return py::make_tuple(v[0], v[1], v[2]); })
        .def_property_readonly("camera_forward",
                               [](const Application& a) { auto v = a.camera_forward(); return py::make_tuple(v[0], v[1], v[2]); })
        // passes
        .def("set_pass_enabled", &Application::set_pass_enabled, py::arg("name"), py::arg("enabled"))
        // tonemap
        .def("set_exposure", &Application::set_exposure, py::arg("exposure"))
        .def_property_readonly("exposure", &Application::exposure)
        // input
        .def("set_input_enabled", &Application::set_input_enabled, py::arg("enabled"))
        .def_property_readonly("input_enabled", &Application::input_enabled)
        // screenshot / window
        .def("capture_screenshot", &Application::capture_screenshot, py::arg("path"))
        .def_property_readonly("window_size",
                               [](const Application& a) { auto v = a.window_size(); return py::make_tuple(v[0], v[1]); })
        // headless additions
        .def("set_window_size", &Application::set_window_size, py::arg("width"), py::arg("height"))
        .def("request_close", &Application::request_close)
        .def_readwrite("max_frames", &Application::max_frames)
        .def_property_readonly("accumulated_spp", &Application::accumulated_spp)
        .def_property_readonly("frames_rendered", &Application::frames_rendered)
        .def("display_image",
             [](const Application& a) {
                 std::vector<float> v = a.display_image();
                 auto ws = a.window_size();
                 py::array_t<float> out({ws[1], ws[0], 4});
                 std::copy(v.begin(), v.end(), out.mutable_data());
                 return out;
             })
        .def("display_ldr",
             [](const Application& a) {
                 std::vector<uint8_t> v = a.display_ldr();
                 auto ws = a.window_size();
                 py::array_t<uint8_t> out({ws[1], ws[0], 4});
                 std::copy(v.begin(), v.end(), out.mutable_data());
                 return out;
             })
        .def_property_readonly("pathtracer_pass", &Application::pathtracer_pass, py::return_value_policy::reference_internal)
        .def_property_readonly("accumulate_pass", &Application::accumulate_pass, py::return_value_policy::reference_internal)
        .def_property_readonly("restir_di_pass", &Application::restir_di_pass, py::return_value_policy::reference_internal)
        .def_property_readonly("restir_pt_pass", &Application::restir_pt_pass, py::return_value_policy::reference_internal)
        .def_property_readonly("restir_gi_pass", &Application::restir_gi_pass, py::return_value_policy::reference_internal)
        .def_property_readonly("atrous_pass", &Application::atrous_pass, py::return_value_policy::reference_internal)
        .def_property_readonly("guiding_pass", &Application::guiding_pass, py::return_value_policy::reference_internal)
        .def_property_readonly("gbuffer_pass", &Application::gbuffer_pass, py::return_value_policy::reference_internal)
        .def_property_readonly("tonemap_pass", &Application::tonemap_pass, py::return_value_policy::reference_internal);

    m.def(
        "load_gltf",
        [](const std::string& path) {
            rrh::scene::GltfScene g;
            {
                py::gil_scoped_release nogil;
                rrh::scene::load_gltf(path, g);
            }
            auto raw = [](const void* p, size_t n) { return py::bytes(static_cast<const char*>(p), n); };
            py::list images;
            for (const auto& im : g.images) {
                py::dict di;
                di["name"] = im.name;
                di["mime_type"] = im.mime_type;
                di["uri"] = im.uri;
                di["width"] = im.width;
                di["height"] = im.height;
                di["bytes"] = im.bytes;
                images.append(di);
            }
            py::list mat_images;
            for (const auto& mi : g.material_images)
                mat_images.append(py::make_tuple(mi.base_color, mi.metallic_roughness, mi.normal, mi.emissive));
            py::dict d;
            d["name"] = g.name;
            d["vertices"] = raw(g.vertices.data(), g.vertices.size() * sizeof(rrh_vertex));
            d["indices"] = raw(g.indices.data(), g.indices.size() * sizeof(uint32_t));
            d["meshes"] = raw(g.meshes.data(), g.meshes.size() * sizeof(rrh_mesh));
            d["materials"] = raw(g.materials.data(), g.materials.size() * sizeof(rrh_material));
            d["instances"] = raw(g.instances.data(), g.instances.size() * sizeof(rrh_instance));
            d["mesh_names"] = g.mesh_names;
            d["material_names"] = g.material_names;
            d["material_images"] = mat_images;
            d["images"] = images;
            d["warnings"] = g.warnings;
            return d;
        },
        py::arg("path"), "Parse a .glb / .gltf file into raw scene records (bytes in the rrh_* layouts).");
    m.def(
        "tonemap_ldr",
        [](py::array_t<float, py::array::c_style | py::array::forcecast> hdr, float exposure) {
            if (hdr.ndim() < 1 || hdr.shape(hdr.ndim() - 1) != 4) throw std::invalid_argument("tonemap_ldr: (..., 4) RGBA");
            const size_t px = (size_t)hdr.size() / 4;
            py::array_t<uint8_t> out(std::vector<py::ssize_t>(hdr.shape(), hdr.shape() + hdr.ndim()));
            rrh::app::tonemap_ldr(hdr.data(), px, exposure, out.mutable_data());
            return out;
        },
        py::arg("hdr"), py::arg("exposure") = 1.0f);
    m.def(
        "write_png",
        [](const std::string& path, py::array_t<uint8_t, py::array::c_style | py::array::forcecast> rgba) {
            if (rgba.ndim() != 3 || rgba.shape(2) != 4) throw std::invalid_argument("write_png: (H, W, 4) uint8");
            if (!rrh::app::write_png(path, rgba.data(), (uint32_t)rgba.shape(1), (uint32_t)rgba.shape(0)))
                throw std::runtime_error("write_png: cannot write " + path);
        },
        py::arg("path"), py::arg("rgba"));
}
