defmodule NxStereo.HipSegment.Records do
  @moduledoc """
  The segment protocol's records, as include/nxs_hip.h takes them and
  native/stereo_match/src/protocol.rs:1-23 defines them (the Python mirror is
  nx_stereo_amd/protocol.py): configure, run and release requests, and the
  replies tagged 7 (configured), 8 (member times and output tensors) and 9
  (refused). Little-endian; a name is its length (u8) and bytes; a tensor is its
  name, type (0 u8, 1 u16, 2 f32, 3 s32, 4 u32), rank, u32 dimensions and
  row-major elements.
  """

  alias NxStereo.Stage

  @configure 1
  @run 2
  @release 3
  @configured 7
  @outputs 8
  @refused 9
  @types %{0 => {:u, 8}, 1 => {:u, 16}, 2 => {:f, 32}, 3 => {:s, 32}, 4 => {:u, 32}}
  @type_codes %{{:u, 8} => 0, {:u, 16} => 1, {:f, 32} => 2, {:s, 32} => 3, {:u, 32} => 4}

  @doc "The configure record of segment `id`: its geometry, then each member's kind and settings."
  def configure(id, members, geometry) do
    [
      <<@configure, id::little-32, geometry.width::little-16, geometry.height::little-16,
        geometry.frame_width::little-16, geometry.frame_height::little-16, length(members)>>,
      Enum.map(members, fn member ->
        settings = settings(member)
        [name(member.module.native_kind()), length(settings),
         Enum.map(settings, fn {key, value} -> [name(key), <<value::little-float-64>>] end)]
      end)
    ]
    |> IO.iodata_to_binary()
  end

  # A member's build settings and the constants its step shares, as f64.
  defp settings(%{module: module, state: state}) do
    build = for {key, value} <- Stage.native_settings(module, state), do: {key, float(value)}

    constants =
      for {key, value} <- module |> Stage.native_constants() |> Enum.sort(),
          do: {key, float(value)}

    build ++ constants
  end

  defp float(true), do: 1.0
  defp float(false), do: 0.0
  defp float(value), do: value * 1.0

  @doc "Each member's params for one pair: their names (nested keys joined by dots) and f32 values."
  def inputs(members, geometry) do
    [
      length(members),
      Enum.map(members, fn member ->
        values = member.module |> Stage.params(member.state, geometry) |> flat("") |> Enum.sort()

        [length(values),
         Enum.map(values, fn {key, binary} ->
           [name(key), <<div(byte_size(binary), 4)::little-16>>, binary]
         end)]
      end)
    ]
    |> IO.iodata_to_binary()
  end

  defp flat(params, prefix) do
    Enum.flat_map(params, fn
      {key, %Nx.Tensor{} = tensor} -> [{prefix <> to_string(key), tensor |> Nx.as_type(:f32) |> Nx.to_binary()}]
      {key, value} when is_number(value) -> [{prefix <> to_string(key), <<value::little-float-32>>}]
      {key, %{} = nested} -> flat(nested, "#{prefix}#{key}.")
    end)
  end

  @doc "The run record of segment `id`: its members' `inputs` (from `inputs/2`) and the entries."
  def run(id, inputs, entries) do
    [<<@run, id::little-32>>, inputs, length(entries), Enum.map(entries, &tensor/1)]
  end

  @doc "The release record of segment `id`."
  def release(id), do: <<@release, id::little-32>>

  defp tensor({field, tensor}) do
    dims = Tuple.to_list(Nx.shape(tensor))

    [name(field), <<Map.fetch!(@type_codes, Nx.type(tensor)), length(dims)>>,
     Enum.map(dims, &<<&1::little-32>>), Nx.to_binary(tensor)]
  end

  defp name(key) do
    key = to_string(key)
    [byte_size(key), key]
  end

  @doc """
  A reply record: `:configured`, `{:outputs, times_us, outputs}` (tensors by name,
  as `Nx.BinaryBackend` tensors), `{:refused, reason}`, `:released` (an empty reply)
  or `{:error, :bad_reply}`.
  """
  def decode(<<>>), do: :released
  def decode(<<@configured>>), do: :configured
  def decode(<<@refused, reason::binary>>), do: {:refused, reason}

  def decode(<<@outputs, count, times::binary-size(count * 4), n, rest::binary>>) do
    times_us = for <<time::little-32 <- times>>, do: time

    case tensors(n, rest, %{}) do
      {:ok, outputs} -> {:outputs, times_us, outputs}
      error -> error
    end
  end

  def decode(_reply), do: {:error, :bad_reply}

  defp tensors(0, <<>>, outputs), do: {:ok, outputs}

  defp tensors(n, <<length, key::binary-size(length), code, rank, rest::binary>>, outputs)
       when is_map_key(@types, code) and byte_size(rest) >= rank * 4 do
    <<dims::binary-size(rank * 4), rest::binary>> = rest
    shape = for <<d::little-32 <- dims>>, do: d
    {_kind, bits} = type = Map.fetch!(@types, code)
    size = Enum.product(shape) * div(bits, 8)

    with <<data::binary-size(size), rest::binary>> <- rest,
         {:ok, name} <- output_name(key) do
      tensor = data |> Nx.from_binary(type, backend: Nx.BinaryBackend) |> Nx.reshape(List.to_tuple(shape))
      tensors(n - 1, rest, Map.put(outputs, name, tensor))
    else
      :error -> {:error, {:unknown_output, key}}
      _short -> {:error, :bad_reply}
    end
  end

  defp tensors(_n, _rest, _outputs), do: {:error, :bad_reply}

  # A name no loaded module mentions is no output of these stages.
  defp output_name(key) do
    {:ok, String.to_existing_atom(key)}
  rescue
    ArgumentError -> :error
  end
end
